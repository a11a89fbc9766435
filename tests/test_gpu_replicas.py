"""Thermal replicas of one cluster: the dipolar field as a product on MFMA.

The product Ht = Mt D^T is checked against the oracle's own pair sum
(field::multi_add_dipolar restated, lib/field.cpp:187-225) replica by
replica: the FP64 product to 1e-12, the FP32 MFMA product to 2e-6 and the
BF16 three-term split to 5e-5 of the field's norm (BF16 keeps 8 mantissa
bits, the split 16; errors add up over 3N = 120 terms).  Whole runs on the
product path are compared with the pair-sum kernel on the same Philox
stream.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")
from magpy_amd import geometry as G  # noqa: E402

MU0 = 1.25663706e-6


def oracle_field(loc, radius, K, ms, m):
    """Reduced dipolar field of one state m [N,3] via the oracle's pair sum."""
    L = O.lib()
    n = len(radius)
    red = O.reduced_parameters(radius, K, ms, 0.1, 0.0, 1.0, 1.0)
    vecs, mags, units = np.zeros(n * n * 3), np.zeros(n * n), np.zeros(n * n * 3)
    L.orc_pair_distance_vectors(O.ptr(vecs), O.ptr(O.f64(loc)), n)
    L.orc_pair_distance_magnitudes(O.ptr(mags), O.ptr(vecs), n)
    L.orc_pair_unit_vectors(O.ptr(units), O.ptr(vecs), O.ptr(mags), n)
    cubes = O.f64((mags / red["v_av"] ** (1. / 3)) ** 3)
    units = O.f64(np.nan_to_num(units))
    h = np.zeros(3 * n)
    v_red = O.f64(red["volume"] / red["v_av"])
    L.orc_field_multi_add_dipolar(O.ptr(h), ms, red["k_av"], O.ptr(v_red), O.ptr(O.f64(m)),
                                  O.ptr(units), O.ptr(cubes), n)
    return h.reshape(n, 3)


def test_dipolar_product_matches_oracle_pair_sum():
    print("CASE dipolar_product", flush=True)
    rng = np.random.default_rng(1)
    n, reps = 40, 130   # padded to 128-multiples on the device
    loc = G.random_cluster(n, 10e-9, 0.15, 2.2, rng)
    radius = 10e-9 * np.exp(0.05 * rng.normal(size=n))
    K = np.full(n, 23e3)
    m = rng.normal(size=(reps, n, 3))
    m /= np.linalg.norm(m, axis=2, keepdims=True)
    ref = np.stack([oracle_field(loc, radius, K, 4e5, m[r]) for r in range(reps)])
    scale = np.linalg.norm(ref, axis=(1, 2), keepdims=True)
    tol = {"pairs": 1e-12, "gemm_f32": 2e-6, "gemm_bf16x3": 5e-5}
    for mode, t in tol.items():
        h = core.dipolar_field(loc, radius, K, 4e5, m, mode=mode)
        err = np.max(np.linalg.norm(h - ref, axis=(1, 2), keepdims=True) / scale)
        print(mode, "max relative error", err, flush=True)
        assert err < t, (mode, err)


def test_512_particle_product_matches_oracle_pair_sum():
    # the shared-geometry workload's cluster: 512 particles at 15 % volume
    # fraction, 3N = 1536 columns, 48 k stages; 128 replicas make 8 column
    # tiles of one row block, so blocks are renumbered for the XCDs
    print("CASE dipolar_product_512", flush=True)
    rng = np.random.default_rng(21)
    n, reps = 512, 128
    loc = G.random_cluster(n, 10e-9, 0.15, 2.2, rng)
    radius = np.full(n, 10e-9)
    K = np.full(n, 23e3)
    m = rng.normal(size=(reps, n, 3))
    m /= np.linalg.norm(m, axis=2, keepdims=True)
    ref = np.stack([oracle_field(loc, radius, K, 4e5, m[r]) for r in range(reps)])
    scale = np.linalg.norm(ref, axis=(1, 2), keepdims=True)
    for mode, t in {"pairs": 1e-12, "gemm_f32": 2e-6, "gemm_bf16x3": 5e-5}.items():
        h = core.dipolar_field(loc, radius, K, 4e5, m, mode=mode)
        err = np.max(np.linalg.norm(h - ref, axis=(1, 2), keepdims=True) / scale)
        print(mode, "max relative error", err, flush=True)
        assert err < t, (mode, err)


def replica_ensemble(n, reps, seed, phi=0.15):
    rng = np.random.default_rng(seed)
    loc = G.random_cluster(n, 10e-9, phi, 2.2, rng)
    axes = G.uniform_random_axes(n, rng)
    P = n * reps
    return dict(member_offsets=np.arange(reps + 1, dtype=np.int64) * n,
                radius=np.full(P, 10e-9), anisotropy=np.full(P, 23e3),
                anisotropy_axis=np.tile(axes, (reps, 1)),
                magnetisation_direction=np.tile(axes, (reps, 1)),
                location=np.tile(loc, (reps, 1)), magnetisation=4e5, damping=0.1,
                temperature=300.0, field_shape="sine", field_amplitude=0.02 / MU0,
                field_frequency=3e8, seeds=rng.integers(0, 2 ** 31 - 1, size=reps))


def run(e, dt, end, S, **kw):
    return core.Session(time_step=dt, end_time=end, max_samples=S, interactions=True,
                        trajectories=True, **e, **kw)


@pytest.mark.parametrize("renorm", [False, True])
def test_replica_path_matches_pair_sums(renorm):
    print("CASE replica_vs_pairs renorm=%s" % renorm, flush=True)
    e = replica_ensemble(32, 64, 2)
    dt, S = 1e-14, 5
    end = 400 * dt
    res = {}
    for mode in ("gemm_f32", "gemm_bf16x3", "pairs"):
        ses = run(e, dt, end, S, precision="fp32", dipolar=mode, renorm=renorm)
        assert ses.info()["dipolar"] == mode
        res[mode] = ses.finish()
    ses = run(e, dt, end, S, precision="fp64", renorm=renorm)
    res["fp64"] = ses.finish()
    for mode, tol in (("gemm_f32", 2e-5), ("gemm_bf16x3", 5e-5), ("pairs", 2e-5)):
        for c in "xyz":
            err = np.max(np.abs(res[mode][c] - res["fp64"][c])) / 4e5
            print(mode, c, err, flush=True)
            assert err < tol, (mode, c, err)
    assert np.all(res["gemm_bf16x3"]["status"] == 0)
    if renorm:
        for mode in ("gemm_f32", "gemm_bf16x3"):
            fs = res[mode]["final_state"]
            assert np.allclose(np.linalg.norm(fs, axis=1), 1.0, atol=2e-6), mode


def test_512_particle_replicas_track_fp64_pair_sums():
    # 200 steps of 128 replicas of the 512-particle cluster on both products
    # against the FP64 pair-sum kernel on the same Philox stream
    print("CASE replica_512_vs_fp64", flush=True)
    e = replica_ensemble(512, 128, 4)
    dt, S = 1e-14, 5
    end = 200 * dt
    res = {}
    for mode in ("gemm_f32", "gemm_bf16x3"):
        ses = run(e, dt, end, S, precision="fp32", dipolar=mode)
        assert ses.info()["dipolar"] == mode
        res[mode] = ses.finish()
    res["fp64"] = run(e, dt, end, S, precision="fp64", dipolar="pairs").finish()
    for mode in ("gemm_f32", "gemm_bf16x3"):
        assert np.all(res[mode]["status"] == 0)
        for c in "xyz":
            err = np.max(np.abs(res[mode][c] - res["fp64"][c])) / 4e5
            print(mode, c, err, flush=True)
            assert err < 1e-5, (mode, c, err)


@pytest.mark.parametrize("renorm", [False, True])
def test_fused_stages_equal_separate_stage_kernels(monkeypatch, renorm):
    # The BF16x3 product runs the Heun stages in its epilogue; with
    # MAGPY_AMD_REPLICA_FUSED=0 it writes the field and the elementwise stage
    # kernels follow.  The field's two k halves are added in either order
    # (commutative) and the stages evaluate the same expressions, so the
    # states are equal bit for bit; the ensemble sums only differ in the
    # order of their per-wave partials.
    print("CASE replica_fused_vs_separate renorm=%s" % renorm, flush=True)
    e = replica_ensemble(512, 128, 5)
    dt, S = 1e-14, 7
    end = 150 * dt
    res = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("MAGPY_AMD_REPLICA_FUSED", fused)
        ses = run(e, dt, end, S, precision="fp32", dipolar="gemm_bf16x3", renorm=renorm)
        res[fused] = ses.finish()
    for c in ("x", "y", "z", "final_state"):
        np.testing.assert_array_equal(res["1"][c], res["0"][c])
    np.testing.assert_allclose(res["1"]["ensemble_m"], res["0"]["ensemble_m"], rtol=1e-12)
    assert np.all(res["1"]["status"] == 0)


def test_auto_selects_the_product_for_replicas():
    print("CASE replica_auto", flush=True)
    e = replica_ensemble(32, 64, 3)
    ses = run(e, 1e-14, 64e-14, 3, precision="fp32")
    assert ses.info()["dipolar"] == "gemm_bf16x3"
    a = ses.finish()
    ses = run(e, 1e-14, 64e-14, 3, precision="fp32")
    b = ses.finish()
    assert np.array_equal(a["final_state"], b["final_state"])
    # different clusters per member: pair sums
    e2 = dict(e)
    e2["location"] = e["location"] * (1 + 1e-3 * np.arange(len(e["radius"]))[:, None])
    ses = run(e2, 1e-14, 64e-14, 3, precision="fp32")
    assert ses.info()["dipolar"] == "pairs"
