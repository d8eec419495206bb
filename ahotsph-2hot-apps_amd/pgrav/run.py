"""Multi-step runs on SDF snapshots: read the outbody table, step it with the
batched force evaluation and the GPU kick-drift, write the result.

    python -m pgrav.run IN.sdf OUT.sdf --steps 10 --dt 0.01 [--theta 0.6 --order 4
        --eps 1e-4 --soft plummer --G 1 --rebuild 1 --variant rel32c]

Each step follows the reference's main loop (nln/main.c:1196-1235): forces at
tpos, then Integrate with dtv = tpos + dt/2 - tvel (which also makes the first
half kick when the snapshot's velocities are synchronised, tvel = tpos) and
dt.  The tree and lists are rebuilt every --rebuild steps (the reference
rebuilds every step); between rebuilds the moments are recomputed on the GPU
for the moved bodies.  The output holds positions at tpos, velocities at tvel,
and the accelerations and potentials at tpos (SAVE_ACC), with ke and pe of the
last step in the header (nln/output.c:270-423).
"""
import argparse
import sys
import time

import numpy as np

from . import batch, integrate, sdf


class Stepper:
    def __init__(self, recs, tpos=0.0, tvel=None, **plan_kw):
        n = recs.shape[0]
        self.t = np.zeros(n, integrate.BODY)
        self.t["mass"] = recs["mass"]
        self.t["pos"] = np.stack([recs["x"], recs["y"], recs["z"]], -1)
        if "vx" in recs.dtype.names:
            self.t["vel"] = np.stack([recs["vx"], recs["vy"], recs["vz"]], -1)
        self.t["ident"] = recs["ident"] if "ident" in recs.dtype.names else np.arange(n)
        self.tpos = float(tpos)
        self.tvel = float(tpos if tvel is None else tvel)
        self.plan_kw = plan_kw
        self.variant = plan_kw.pop("variant", "rel32c")
        self.rebuild = max(1, int(plan_kw.pop("rebuild", 1)))
        self.plan = None
        self.nforce = 0
        self.ke = self.pe = 0.0
        self.self_phi = 0.0
        if plan_kw.get("soft", "plummer") == "plummer" and plan_kw.get("eps", 1e-4) > 0:
            # the softened self pair each sink keeps: -G m / eps
            self.self_phi = plan_kw.get("G", 1.0) / plan_kw.get("eps", 1e-4)

    def forces(self):
        pos = self.t["pos"].astype(np.float64)
        mass = self.t["mass"].astype(np.float64)
        if self.plan is None or self.nforce % self.rebuild == 0:
            if self.plan is not None:
                self.plan.close()
            self.plan = batch.Plan(pos, mass, **self.plan_kw)
        else:
            self.plan.update(pos, mass)
        self.plan.evaluate(self.variant)
        r = self.plan.result(self.variant)
        self.t["acc"] = r[:, :3]
        self.t["phi"] = r[:, 3] + self.self_phi * mass
        self.nforce += 1

    def step(self, dt):
        self.forces()
        dtv = self.tpos + 0.5 * dt - self.tvel  # nln/main.c:1229
        self.tpos, self.tvel, self.ke, self.pe = integrate.integrate(self.t, dt, dtv, self.tpos,
                                                                     self.tvel)

    def records(self):
        out = np.zeros(self.t.shape[0], sdf.OUTBODY_ACC)
        out["mass"] = self.t["mass"]
        for k, c in enumerate("xyz"):
            out[c] = self.t["pos"][:, k]
            out["v" + c] = self.t["vel"][:, k]
            out["a" + c] = self.t["acc"][:, k]
        out["phi"] = self.t["phi"]
        out["ident"] = self.t["ident"]
        return out

    def close(self):
        if self.plan is not None:
            self.plan.close()
            self.plan = None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("input")
    ap.add_argument("output")
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--dt", type=float, required=True)
    ap.add_argument("--theta", type=float, default=0.6)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--soft", default="plummer", choices=sorted(batch.SOFT))
    ap.add_argument("--G", type=float, default=1.0)
    ap.add_argument("--rebuild", type=int, default=1)
    ap.add_argument("--variant", default="rel32c")
    a = ap.parse_args(argv)
    import torch  # noqa: F401  (one HIP runtime in the process: torch's)
    hdr, recs = sdf.read(a.input)
    st = Stepper(recs, tpos=hdr.get("tpos", 0.0), tvel=hdr.get("tvel"), theta=a.theta,
                 order=a.order, eps=a.eps, soft=a.soft, G=a.G, rebuild=a.rebuild,
                 variant=a.variant)
    t0 = time.perf_counter()
    for s in range(a.steps):
        st.step(a.dt)
        print("step %d: tpos %.6g ke %.9g pe %.9g E %.9g" % (s + 1, st.tpos, st.ke, st.pe,
                                                             st.ke + st.pe),
              file=sys.stderr, flush=True)
    st.forces()  # accelerations at the final positions
    wall = time.perf_counter() - t0
    iter_ = int(hdr.get("iter", 0)) + a.steps
    out_hdr = sdf.output_header(recs.shape[0], iter_=iter_, tpos=st.tpos, tvel=st.tvel,
                                tacc=st.tpos, eps=a.eps, tol=a.theta, ke=st.ke, pe=st.pe, G=a.G)
    sdf.write(a.output, st.records(), sdf.OUTBODY_DESC_ACC, out_hdr)
    st.close()
    print("%d steps of %d bodies in %.2f s" % (a.steps, recs.shape[0], wall), file=sys.stderr)


if __name__ == "__main__":
    main()
