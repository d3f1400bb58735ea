"""The benchmark configurations at their real sizes, on the GPU, each against its oracle.

* Single sinogram N=256 in FP64 against the reference's own engine (cpp/oblique.cpp row kernels under the oracle's
  host) at dense_cap p in {8, 16, 32}, plus the FP32 plan against that oracle.
* The 512-sinogram uneven batch (N from {64, 128, 256, 512}), FP32 in one graph launch, each sinogram against its FP64
  plan, and a sinogram alone equal to itself in the batch.
* The serving loop at N in {128, 256, 512, 1024} and p in {16, 32}: every result equal, bit for bit, to the same
  request run alone, and within 1e-5 of FP64.
* N=2048 at batch 2: the FP32 plan against the FP64 plan.
"""

import os

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _sino(n, seed, noise_seed, level=0.01):
    ph = pkg("phantom")
    R = pkg("radon")
    return ph.add_noise(R.adrt(ph.ellipses(seed, 1 << n)), level, noise_seed)


def _oracle_hybrid(n, H, cap):
    """The oracle's FastHybrid on the reference's C++ row kernels (oracle/_ref) when built, else on its NumPy rows,
    with the GPU plan's setup scalars."""
    from oracle import inverses as oiv
    from oracle import refcpp

    threads = os.cpu_count() or 1
    if refcpp.available():
        class Ref(oiv.Hybrid):
            engine_cls = refcpp.RefRows

        return Ref(n, lam_max=H.lam_max, dense_cap=cap, threads=threads, outer=(H.lo, H.hi, H.outer_deg))
    return oiv.Hybrid(n, lam_max=H.lam_max, dense_cap=cap, outer=(H.lo, H.hi, H.outer_deg))


@pytest.mark.parametrize("cap", [8, 16, 32])
def test_n256_single_sinogram_fp64_vs_reference_engine(gpu, cap):
    """N=256, ellipse phantom seed 1 + 1% noise seed 2, dense_cap p: the FP64 GPU hybrid's G_h, backprojection and
    whole inverse within 1e-12 of the reference engine on the same setup scalars; the FP32 plan (TwoSum rows, one
    FP64 refinement step) within 1e-5 of it, at max_batch 1 and at the headline's 128 (the paired inverse FFT)."""
    I = pkg("inverses")
    ph = pkg("phantom")
    n = 8
    b = _sino(n, 1, 2)
    x = ph.ellipses(1, 1 << n)
    H = I.FastHybridHip(n, dense_cap=cap)
    assert H.eng.n_dense == int(np.sum(H.deg > cap)) > 0
    R = _oracle_hybrid(n, H, cap)
    assert np.array_equal(R.deg, H.deg) and R.eng.n_dense == H.eng.n_dense
    assert relerr(H.Gh(x), R.Gh(x)) < 1e-12
    assert relerr(H.backproject(b), R.backproject(b)) < 1e-12
    ref = R.apply(b)
    assert relerr(H.apply(b), ref) < 1e-12
    H32 = I.FastHybridHip(n, lam_max=H.lam_max, dense_cap=cap, precision="fp32", refine=1,
                          outer=(H.lo, H.hi, H.outer_deg))
    assert relerr(H32.apply(b), ref) < 1e-5
    # the headline's plan shape: max_batch 128 gives its FP32 passes the paired inverse FFT
    H32p = I.FastHybridHip(n, lam_max=H.lam_max, dense_cap=cap, precision="fp32", refine=1, max_batch=128,
                           outer=(H.lo, H.hi, H.outer_deg))
    assert relerr(H32p.apply(b), ref) < 1e-5
    H32p.close()
    H32.close()
    H.close()


def test_mixed_batch_512_sinograms(gpu):
    """The uneven batch of the benchmark: 512 sinograms, N drawn from {64: .4, 128: .3, 256: .2, 512: .1}
    (default_rng(42)), phantoms 1000+i, 1% noise 2000+i, FP32 hybrid (dense_cap 16, one FP64 refinement step) in one
    graph launch.  Every sinogram within 1e-5 of its FP64 plan; one sinogram of each size alone gives the bits it
    gives in the batch."""
    B = pkg("batch")
    I = pkg("inverses")
    C = pkg("plancache")
    ns = B.draw_sizes(512)
    sinos = [_sino(n, 1000 + i, 2000 + i) for i, n in enumerate(ns)]
    mb = B.MixedBatch(ns, precision="fp32", refine=1, dense_cap=16)
    mb.set_inputs(sinos)
    mb.run()
    x = mb.outputs()
    errs = []
    for n, idx in sorted(mb.groups.items()):
        lam, outer = C.hybrid_setup(n, dense_cap=16)
        H64 = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=16, max_batch=len(idx))
        x64 = H64.apply_batch(np.stack([sinos[i] for i in idx]))
        for j, i in enumerate(idx):
            errs.append(relerr(x[i], x64[j]))
            assert errs[-1] < 1e-5, (i, n, errs[-1])
        H64.close()
        i = idx[len(idx) // 2]
        assert np.array_equal(mb.plans[n].apply(sinos[i]), x[i]), (i, n)
    assert len(errs) == 512
    mb.close()


def test_serving_n128_to_1024(gpu):
    """The serving workload's sizes and caps (N in {128, 256, 512, 1024}, p in {16, 32}), each FP32 plan with the
    FP64 refinement steps refine_for picks: requests released over a short trace all complete; each result equals,
    bit for bit, the same request run alone through its slot's plan, and is within 1e-5 of the FP64 inverse with the
    same setup."""
    S = pkg("serve")
    C = pkg("plancache")
    I = pkg("inverses")
    A = pkg("autotune")
    keys = [(n, p) for n in (7, 8, 9, 10) for p in (16, 32)]
    setups = {k: C.hybrid_setup(k[0], dense_cap=k[1]) for k in keys}
    refine = {k: A.refine_for(k[0], k[1], *setups[k]) for k in keys}
    assert refine[(7, 32)] > refine[(7, 16)]  # no dense rows at N=128, p=32: more FP64 steps for the same error
    srv = S.Server(keys, slots_per_key=1, max_batch={7: 4, 8: 4, 9: 2, 10: 2}, setups=setups, refine=refine)
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(24):
        n, p = keys[i % len(keys)] if i < len(keys) else keys[int(rng.integers(len(keys)))]
        reqs.append(S.Request(i, n, p, _sino(n, 900 + i, 950 + i), 0.0))
    trace = [(0.002 * i, r) for i, r in enumerate(reqs)]
    done = srv.run_trace(trace)
    assert len(done) == len(reqs) and all(r.error is None for r in reqs)
    for r in reqs:
        slot = next(s for s in srv.slots if (s.n, s.cap) == (r.n, r.cap))
        assert np.array_equal(r.x, slot.plan.apply(r.b)), r.rid
    for n, p in keys:  # one request per key against FP64
        r = next(q for q in reqs if (q.n, q.cap) == (n, p))
        lam, outer = setups[(n, p)]
        H64 = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=p)
        assert relerr(r.x, H64.apply(r.b)) < 1e-5, (n, p)
        H64.close()
    srv.close()


def test_n2048_fp32_vs_fp64_batch2(gpu):
    """N=2048, dense_cap 16, ellipse phantoms + 0.5% noise, batch 2: the FP32 plan (TwoSum rows, FP64 dense rows,
    outer iteration and refinement) within 1e-5 of the FP64 plan for both sinograms, and a sinogram alone equal to
    itself in the batch."""
    I = pkg("inverses")
    n = 11
    Bn = np.stack([_sino(n, 5000 + i, 6000 + i, level=0.005) for i in range(2)])
    H64 = I.FastHybridHip(n, dense_cap=16, max_batch=2)
    x64 = H64.apply_batch(Bn)
    H32 = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=16, precision="fp32", refine=1, max_batch=2,
                          outer=(H64.lo, H64.hi, H64.outer_deg))
    x32 = H32.apply_batch(Bn)
    for i in range(2):
        assert relerr(x32[i], x64[i]) < 1e-5, i
    assert np.array_equal(H32.apply(Bn[1]), x32[1])
    H32.close()
    H64.close()
