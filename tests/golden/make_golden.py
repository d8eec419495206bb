"""Writes tests/golden/ref_kernels.npz: input/output vectors of the reference's
own CPU kernels, compiled from the reference sources into oracle/_ref/ by
`make -C oracle ref`.

    python tests/golden/make_golden.py

Kernels (all from lanl/ahotsph-2hot-apps nln/):
  pHinteract, pQinteract, pMinteract, pMinteract_sK1    pgrav_vec.c:45-401 (NSSE 8)
  do_gravph_sse4, do_gravpq_sse4, do_gravp_sse4          do_grav_sse4.c:1768-2224
  _pgravm_sK1                                            py_pgrav/_pgrav.c:10-49
on seeded inputs built here.  Each vector is what the reference code returned;
tests/test_oracle.py checks the FP64 oracle against them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def cells(rng, ncell, nbody=40, spread=0.08, box=1.0):
    """Seeded order-4 cell records (nln/cofm.c:398-562 moments about the cofm)."""
    recs = []
    for _ in range(ncell):
        c0 = rng.uniform(-box, box, 3)
        pos = c0 + rng.normal(size=(nbody, 3)) * spread * rng.uniform(0.3, 1.0, 3)
        mass = rng.uniform(0.5, 1.5, nbody) / nbody
        cm = (pos * mass[:, None]).sum(0) / mass.sum()
        bmax = np.sqrt(((pos - cm) ** 2).sum(1)).max()
        rec = oracle.cell_record(pos, mass, cm, bmax)
        # non-zero dipole so that the dipole path of pH/pQ is exercised
        rec[5:8] = rng.normal(size=3) * 0.05 * rec[0]
        recs.append(rec)
    return np.array(recs)


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref not built: make -C oracle ref")
    rng = np.random.default_rng(20261202)
    out = {}
    ncell, m = 40, 24
    recs = cells(rng, ncell)
    sinks = rng.uniform(-3.0, 3.0, (m, 3)).astype(np.float32)
    body = np.zeros((m, 14), np.float32)
    body[:, 0] = 1.0 / m
    body[:, 1:4] = sinks
    out["cells"] = recs.astype(np.float32)
    out["sinks"] = sinks
    nb = (ncell + 7) // 8
    for kind, sz in (("H", 29), ("Q", 13), ("M", 4)):
        f = oracle.to_soa(recs[:, :sz], 8)
        out["vec_" + kind] = oracle.ref_mn(kind, body, f, nb)
    # K1 monopoles: place some sources inside the softening radius
    eps = 0.8
    near = sinks[:8] + rng.normal(size=(8, 3)).astype(np.float32) * 0.3
    mono = np.concatenate([recs[:, :4].astype(np.float32),
                           np.c_[np.full(8, 0.02, np.float32), near]])
    out["mono_k1"] = mono
    out["vec_MK1"] = oracle.ref_mn("MK1", body, oracle.to_soa(mono, 8), (mono.shape[0] + 7) // 8,
                                   eps=1.0 / eps)
    out["eps_k1"] = np.float64(eps)
    # Plummer-softened SSE4 twins, one sink at a time
    eps2 = 0.01
    out["eps2"] = np.float64(eps2)
    for name, sz in (("do_gravph", 29), ("do_gravpq", 13), ("do_gravp", 4)):
        f4 = oracle.to_soa(recs[:, :sz] if sz != 4 else mono, 4)
        n = ncell if sz != 4 else mono.shape[0]
        out["sse4_" + name] = np.array([oracle.ref_do_grav(name, "sse4", f4, n, s, eps2)
                                        for s in sinks])
    # nln/py_pgrav/test_pgrav.py:6-21 inputs (np.random.seed(123)), K1 version
    rs = np.random.RandomState(123)
    n = 16
    src = np.empty((n, 4), np.float32)
    src[:, 0] = 1.0 / n
    src[:, 1] = rs.uniform(0.0, 1.0, n)
    src[:, 2] = rs.uniform(-0.5, 0.5, n)
    src[:, 3] = rs.uniform(-0.5, 0.5, n)
    sink = np.array([[1.0, 0.0, 0.0, 0.0], [1.0, 0.5, 0.0, 0.1]], np.float32)
    out["tp_src"] = src
    out["tp_sink"] = sink
    out["tp_pgravm_sK1"] = oracle.ref_pgravm_sk1(sink, src, 1.0 / 0.3)
    np.savez_compressed(os.path.join(HERE, "ref_kernels.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_kernels.npz"), sorted(out))


if __name__ == "__main__":
    main()
