"""simulate.py -- the reference's time loop (utils/vvflow/vvflow.cpp:177-266)
on one device context, for a rigid static body in a uniform stream.

One step, in the reference's order:
    tree build; calc_circulation (body SLAE, FP64); tree destroy;
    heat_shed; vortex_shed; streak_shed; calc_forces; (stepdata);
    zero_forces; tree build; epsilon (merge when viscous); convective;
    diffusive (vortices, then heat); tree destroy;
    move_and_clean(remove = True); heat_crop; time += dt.
Tree parameters follow vvflow.cpp:200-203: one tree object (far criterion
8 by default) for both builds, minimum node size 5 dl and maximum 100 dl with
dl the body's average segment length.

The convective step can use the learned far field instead of the plain
tree path (``learned=True``; a model must be loaded into the context).
"""
import time

import numpy as np

from . import BODY_DTYPE, SEG_DTYPE, Context, gen_cylinder, segments_from_corners


class CylinderRun:
    """Impulsively started circular cylinder (gen_cylinder{R, N}) in a
    stream U = (1, 0) at Reynolds number re, time step dt."""

    def __init__(self, ctx=None, radius=0.5, nseg=500, re=600.0, dt=0.005, far_criteria=8, learned=False,
                 kahan=True, heat=None, streak=None):
        """heat: (condition, heat_const, pr) of the body (HC_* condition,
        TBody::heat_condition); streak: (n, dt_streak), n streak sources on
        x = -1, y in [-1, 1] shed every dt_streak."""
        self.ctx = ctx if ctx is not None else Context(0)
        self.re, self.dt, self.fc = float(re), float(dt), int(far_criteria)
        self.learned, self.kahan = learned, kahan
        self.inf = (1.0, 0.0)
        corners = gen_cylinder(radius, nseg)
        segs = segments_from_corners(corners)
        body = np.zeros(1, dtype=BODY_DTYPE)
        body["nseg"] = len(segs)
        self.ctx.set_params(re=self.re, dt=self.dt, time=0.0)
        self.ctx.upload_bodies(np.ascontiguousarray(segs, dtype=SEG_DTYPE), body)
        self.ctx.upload(np.zeros((0, 6)))
        self.heat = heat
        if heat is not None:
            cond, hc, pr = heat
            self.ctx.set_heat([cond], np.full(len(segs), hc, np.float32))
            self.ctx.set_pr(pr)
        self.dt_streak = 0.0
        if streak is not None:
            n, self.dt_streak = streak
            src = np.zeros((n, 6))
            src[:, 0] = -1.0
            src[:, 1] = np.linspace(-1.0, 1.0, n) if n > 1 else 0.0
            self.ctx.upload_list(3, src)   # LIST_STREAK_SOURCE
        dl = self.ctx.download_bodies()[1][0]["slen"] / (nseg - 1)   # Space::average_segment_length
        self.min_node, self.max_node = 5 * dl, 100 * dl
        self.step_no = 0
        self.time = 0.0
        self.forces = []
        self.phase_ms = np.zeros(8)
        # velocity probes (utils/vvflow/sensors.cpp): points whose convective
        # velocity is recorded every step right after the convective call
        self.sensor_points = None
        self.sensor_rows = []

    def step(self):
        c = self.ctx
        c.set_params(re=self.re, dt=self.dt, time=self.time)
        c.tree_build(self.fc, self.min_node, self.max_node)
        c.calc_circulation(self.inf)
        c.tree_destroy()
        c.heat_shed()
        c.vortex_shed()
        if self.dt_streak:
            k = self.time / self.dt_streak   # TTime::divisibleBy on the dt grid
            if abs(k - round(k)) <= 1e-9 * max(1.0, abs(k)):
                c.streak_shed()
        c.calc_forces()
        body = c.download_bodies()[1][0]
        self.forces.append(np.concatenate([body["force_hydro"], body["friction"]]))
        c.zero_forces()
        c.tree_build(self.fc, self.min_node, self.max_node)
        c.epsilon(True)
        if self.learned:
            c.convective_learned(self.inf, kahan=self.kahan)
        else:
            c.convective(self.inf, kahan=self.kahan)
        if self.sensor_points is not None:
            self.sensor_rows.append(c.velocity(self.sensor_points, self.inf))
        c.diffusive()
        c.diffusive_heat()
        c.tree_destroy()
        c.move_and_clean(True)
        c.heat_crop(16.0)
        self.phase_ms += np.maximum(c.phase_times(), 0)
        self.step_no += 1
        self.time = self.step_no * self.dt   # TTime::add keeps time on the dt grid

    def run(self, steps, keep_every=0, progress=None):
        """Runs `steps` steps; returns the vortex lists kept every
        `keep_every` steps (after move_and_clean)."""
        kept = []
        t0 = time.perf_counter()
        for k in range(steps):
            self.step()
            if keep_every and (k + 1) % keep_every == 0:
                kept.append(self.ctx.download())
            if progress and (k + 1) % progress == 0:
                print(f"  re={self.re:g} step {k + 1}/{steps} n={self.ctx.n} "
                      f"({time.perf_counter() - t0:.1f} s)", flush=True)
        return kept
