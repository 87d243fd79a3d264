"""Stage-by-stage parity of the HIP kernels with the oracle on identical
inputs, through the C ABI (the reference-named Python mirror in
demso_amd.api).  Integer outputs (hash, index, permutation, cellStart,
cellEnd, loc) must match bit for bit; FP32 outputs must agree within 1e-5
relative (max |gpu-oracle| / max |oracle| per array) -- both sides compute the
same IEEE sequence without FMA contraction, so they are expected to match
exactly and the tests also report that.
"""
import numpy as np
import pytest

from demso_amd import api
from demso_amd import scene as sc
from helpers import bits_equal, compressed_pack, dev, host, rel_err, state_dict

pytestmark = pytest.mark.gpu
TOL = 1e-5


def scenes():
    s_pack = compressed_pack()
    s_settle = sc.settling(side=128)
    s_pile = sc.triangle_pile()
    s_ens = sc.ensemble(instances=6, counts=[700, 1200, 300, 64, 2500, 900], seed=3)
    # squeeze the ensemble too, so that contacts (and per-instance materials) matter
    for k in range(6):
        a, b = s_ens.segments[k], s_ens.segments[k + 1]
        s_ens.pos[a:b] *= np.float32(0.86)
    s_ens.vel[:] = np.random.default_rng(0).normal(size=s_ens.vel.shape).astype(np.float32) * 0.1
    s_ens.omega[:] = np.random.default_rng(1).normal(size=s_ens.n).astype(np.float32) * 10
    return {"pack": s_pack, "settling128": s_settle, "pile_wrap_attract": s_pile, "ensemble6": s_ens}


SCENES = scenes()


def oracle_sorted(oracle, s):
    system = oracle.System(s)
    h, i = oracle.calc_hash(system, s.pos, s.type)
    hs, is_ = oracle.sort_particles(h, i)
    cs, ce, st = oracle.reorder(hs, is_, state_dict(s), s.total_cells)
    return system, h, i, hs, is_, cs, ce, st


PER = [1, 2, 4]  # particles a lane of hash, reorder and integrate


@pytest.mark.parametrize("per", PER)
@pytest.mark.parametrize("name", list(SCENES))
def test_hash(gpu_ctx, oracle, name, per):
    s = SCENES[name]
    gpu_ctx.load_scene(s)
    gpu_ctx.set_step_variant(per_lane={"hash": per, "reorder": per, "integrate": per})
    system = oracle.System(s)
    h_ref, i_ref = oracle.calc_hash(system, s.pos, s.type)
    pos, typ = dev(s.pos), dev(s.type)
    h, i = dev(np.zeros(s.n, np.uint32)), dev(np.zeros(s.n, np.uint32))
    api.calcHash(gpu_ctx, pos, i, h, s.n, type=typ if len(s.instances) > 1 else None)
    gpu_ctx.synchronize()
    assert np.array_equal(host(h, np.uint32), h_ref)
    assert np.array_equal(host(i, np.uint32), i_ref)


@pytest.mark.parametrize("flags", [1, 0])
def test_hash_edge_positions(gpu_ctx, oracle, flags):
    """Boundaries, out-of-box, negative, huge, inf and NaN coordinates."""
    from test_oracle_kat import one_type_scene

    vals = [0.0, -0.0, 0.5, 7.999999, 8.0, 8.000001, -1e-30, -0.25, 1e9, -1e9, 3e38, np.inf, -np.inf, np.nan,
            1e-45, 4.0, 3.9999998]
    xs = np.array(vals, np.float32)
    pos = np.stack(np.meshgrid(xs, xs), -1).reshape(-1, 2).astype(np.float32)
    s = one_type_scene(pos, box=(8.0, 8.0), grid=(16, 16), flags=flags)
    gpu_ctx.load_scene(s)
    h_ref, _ = oracle.calc_hash(oracle.System(s), s.pos)
    h, i = dev(np.zeros(s.n, np.uint32)), dev(np.zeros(s.n, np.uint32))
    api.calcHash(gpu_ctx, dev(s.pos), i, h, s.n)
    gpu_ctx.synchronize()
    assert np.array_equal(host(h, np.uint32), h_ref)


def _sort_case(ctx, keys, vals):
    n = keys.shape[0]
    h, v = dev(keys), dev(vals)
    api.sortParticles(ctx, h, v, n)
    ctx.synchronize()
    order = np.argsort(keys, kind="stable")
    return np.array_equal(host(h, np.uint32), keys[order]) and np.array_equal(host(v, np.uint32), vals[order])


@pytest.mark.parametrize("n", [2, 63, 64, 65, 4095, 4096, 4097, 65549, 1000003])
@pytest.mark.parametrize("cells", [200, 4186, 11193546])
def test_sort_stable_bounded_keys(gpu_ctx, n, cells):
    """Keys below the configured cell count (2 or 3 radix passes)."""
    s = sc.column_collapse(cols=4, rows=4)
    s.instances[0] = sc.Instance(box=(1.0, 1.0), extent=(1.0, 1.0), grid=(cells, 1))
    gpu_ctx.load_scene(s)
    rng = np.random.default_rng(n + cells)
    keys = rng.integers(0, min(cells, max(2, n // 3)), size=n).astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    assert _sort_case(gpu_ctx, keys, vals)
    # nearly sorted keys: the step's real input (previous step's order)
    keys = np.sort(rng.integers(0, cells, size=n)).astype(np.uint32)
    swap = rng.integers(0, max(1, n - 1), size=max(1, n // 50))
    keys[swap], keys[np.minimum(swap + 1, n - 1)] = keys[np.minimum(swap + 1, n - 1)], keys[swap]
    assert _sort_case(gpu_ctx, keys, np.arange(n, dtype=np.uint32))


def _digit_patterns(n, rng):
    """Keys whose 64-key groups share digits in every arrangement the ranking
    distinguishes: one digit a group, two contiguous halves, interleaved
    repeats, sorted runs, descending keys, and random keys whose digits
    take only two values each."""
    i = np.arange(n, dtype=np.int64)
    two = (rng.integers(0, 2, size=(n, 3)) * np.array([0x80, 0x8000, 0x800000])).sum(1)
    return {
        "constant": np.full(n, 0x123456, np.int64),
        "halves": np.where(i % 64 < 32, 0x00A1B2, 0x00A1B3),
        "interleaved": 0x10101 * (i % 3),
        "runs": i // 7,
        "descending": n - 1 - i,
        "two_values_a_digit": two,
    }


@pytest.mark.parametrize("n", [64, 4097, 65549])
def test_sort_digit_patterns(gpu_ctx, n):
    """Stable order for key layouts that take the one-digit-a-group shortcut,
    the LDS peer masks, or both within one tile (24-bit keys, 3 passes)."""
    s = sc.column_collapse(cols=4, rows=4)
    s.instances[0] = sc.Instance(box=(1.0, 1.0), extent=(1.0, 1.0), grid=(11193546, 1))
    gpu_ctx.load_scene(s)
    rng = np.random.default_rng(n)
    for name, k in _digit_patterns(n, rng).items():
        keys = (k % 11193546).astype(np.uint32)
        assert _sort_case(gpu_ctx, keys, rng.permutation(n).astype(np.uint32)), name


@pytest.mark.parametrize("n", [2, 4097, 65549, 300007])
@pytest.mark.parametrize("shift", [(1, 0), (0, 3), (2, 1)], ids=["keys+4B", "vals+12B", "both"])
def test_sort_unaligned_arrays(gpu_ctx, n, shift):
    """Caller arrays that are not 16-byte aligned (views at an offset of 1-3
    words) take the sort's one-word loads: the same stable order."""
    s = sc.column_collapse(cols=4, rows=4)
    s.instances[0] = sc.Instance(box=(1.0, 1.0), extent=(1.0, 1.0), grid=(11193546, 1))
    gpu_ctx.load_scene(s)
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 11193546, size=n).astype(np.uint32)
    keys[: n // 3] = np.sort(keys[: n // 3])  # a nearly sorted stretch too
    vals = rng.permutation(n).astype(np.uint32)
    dk, dv = shift
    hk = dev(np.concatenate([np.zeros(dk, np.uint32), keys]))
    hv = dev(np.concatenate([np.zeros(dv, np.uint32), vals]))
    kview, vview = hk[dk:], hv[dv:]
    assert (kview.data_ptr() % 16 != 0) or (vview.data_ptr() % 16 != 0)
    api.sortParticles(gpu_ctx, kview, vview, n)
    gpu_ctx.synchronize()
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(host(kview, np.uint32), keys[order])
    assert np.array_equal(host(vview, np.uint32), vals[order])
    assert np.array_equal(host(hk, np.uint32)[:dk], np.zeros(dk, np.uint32))  # nothing written before the views
    assert np.array_equal(host(hv, np.uint32)[:dv], np.zeros(dv, np.uint32))


@pytest.mark.parametrize("n", [0, 1, 3, 5000, 300007])
def test_sort_full_32bit_keys(n):
    """Without parameters the sort covers all 32 key bits (4 passes)."""
    ctx = api.Context(0)
    try:
        rng = np.random.default_rng(n)
        keys = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 4] = 0xFFFFFFFF  # ties at the largest key
        vals = np.arange(n, dtype=np.uint32)
        if n == 0:
            api.sortParticles(ctx, dev(np.zeros(1, np.uint32)), dev(np.zeros(1, np.uint32)), 0)
        else:
            assert _sort_case(ctx, keys, vals)
    finally:
        ctx.close()


@pytest.mark.parametrize("per", PER)
@pytest.mark.parametrize("name", list(SCENES))
def test_reorder(gpu_ctx, oracle, name, per):
    s = SCENES[name]
    gpu_ctx.load_scene(s)
    gpu_ctx.set_step_variant(per_lane={"hash": per, "reorder": per, "integrate": per})
    system, h, i, hs, is_, cs_ref, ce_ref, st_ref = oracle_sorted(oracle, s)
    n, C = s.n, s.total_cells
    rad = s.radius is not None
    out = api.DeviceState(n, rad)
    cs, ce = dev(np.zeros(C, np.uint32)), dev(np.zeros(C, np.uint32))
    api.reorderDataAndFindCellStart(
        gpu_ctx, cs, ce, out.pos, out.vel, out.theta, out.omega, out.id, out.loc, out.type, dev(hs), dev(is_),
        dev(s.pos), dev(s.vel), dev(s.theta), dev(s.omega), dev(s.id), dev(s.type), n, C,
        sortedRadius=out.radius, oldRadius=dev(s.radius) if rad else None)
    gpu_ctx.synchronize()
    cs_g, ce_g = host(cs, np.uint32), host(ce, np.uint32)
    assert np.array_equal(cs_g, cs_ref)
    occ = cs_ref != 0xFFFFFFFF
    assert np.array_equal(ce_g[occ], ce_ref[occ])
    for k, like in (("pos", np.float32), ("vel", np.float32), ("theta", np.float32), ("omega", np.float32),
                    ("id", np.uint32), ("type", np.uint32), ("loc", np.uint32)):
        assert bits_equal(host(getattr(out, k), like), st_ref[k]), k
    if rad:
        assert bits_equal(host(out.radius), st_ref["radius"])


COLLIDE_VARIANTS = {"plain": {}, "prefetch": {"prefetch": True}, "batch": {"batch": True}, "wide": {"wide": True},
                    "list": {"contact_lists": True}}


@pytest.mark.parametrize("variant", list(COLLIDE_VARIANTS))
@pytest.mark.parametrize("name", list(SCENES))
def test_collide(gpu_ctx, oracle, name, variant):
    """The drop-in collide stage in each candidate-loading mode and with
    64-bit addressing: acc, alpha and pressure equal the oracle's bit for bit."""
    s = SCENES[name]
    gpu_ctx.load_scene(s)
    system, h, i, hs, is_, cs, ce, st = oracle_sorted(oracle, s)
    acc_ref, alpha_ref, pres_ref = oracle.collide(system, st, cs, ce)
    assert np.count_nonzero(pres_ref) > 0 or name == "settling128"
    n, C = s.n, s.total_cells
    acc, alpha, pres = dev(np.zeros((n, 2), np.float32)), dev(np.zeros(n, np.float32)), dev(np.full(n, 7.0, np.float32))
    gpu_ctx.set_step_variant(**COLLIDE_VARIANTS[variant])
    try:
        api.collide(gpu_ctx, dev(st["pos"]), dev(st["vel"]), acc, dev(st["omega"]), alpha, dev(st["type"]), dev(cs),
                    dev(ce), n, C, pres, radius=dev(st["radius"]) if st["radius"] is not None else None)
        gpu_ctx.synchronize()
    finally:
        gpu_ctx.set_step_variant()
    a, al, p = host(acc), host(alpha), host(pres)
    assert bits_equal(a, acc_ref), rel_err(a, acc_ref)
    assert bits_equal(al, alpha_ref), rel_err(al, alpha_ref)
    assert bits_equal(p, pres_ref), rel_err(p, pres_ref)


def integrate_inputs(s, seed=0):
    """A state with discs beyond every wall plus random accelerations."""
    rng = np.random.default_rng(seed)
    st = {k: (v.copy() if v is not None else None) for k, v in state_dict(s).items()}
    n = s.n
    r = st["radius"] if st["radius"] is not None else np.full(n, s.materials[0].radius, np.float32)
    box = s.instances[0].box
    q = n // 5
    st["pos"][:q, 1] = (r[:q] * rng.uniform(0.2, 1.05, q)).astype(np.float32)       # bottom wall
    st["pos"][q:2 * q, 0] = (r[q:2 * q] * rng.uniform(0.2, 1.05, q)).astype(np.float32)  # left wall
    st["pos"][2 * q:3 * q, 0] = (box[0] - r[2 * q:3 * q] * rng.uniform(0.2, 1.05, q)).astype(np.float32)
    st["pos"][3 * q:4 * q, 1] = (box[1] - r[3 * q:4 * q] * rng.uniform(0.2, 1.05, q)).astype(np.float32)
    st["vel"] = (rng.normal(size=(n, 2)) * 0.5).astype(np.float32)
    st["omega"] = (rng.normal(size=n) * 50).astype(np.float32)
    st["theta"] = rng.uniform(-7, 13, n).astype(np.float32)
    acc = (rng.normal(size=(n, 2)) * 20).astype(np.float32)
    alpha = (rng.normal(size=n) * 1e3).astype(np.float32)
    pres = np.abs(rng.normal(size=n)).astype(np.float32)
    return st, acc, alpha, pres


@pytest.mark.parametrize("per", PER)
@pytest.mark.parametrize("name", list(SCENES))
def test_integrate(gpu_ctx, oracle, name, per):
    s = SCENES[name]
    gpu_ctx.load_scene(s)
    gpu_ctx.set_step_variant(per_lane={"hash": per, "reorder": per, "integrate": per})
    system = oracle.System(s)
    st, acc, alpha, pres = integrate_inputs(s)
    out_ref, pres_ref = oracle.integrate(system, st, acc, alpha, pres)
    pos, vel, th, om, pr = dev(st["pos"]), dev(st["vel"]), dev(st["theta"]), dev(st["omega"]), dev(pres)
    api.integrateSystem(gpu_ctx, pos, vel, dev(acc), th, om, dev(alpha), dev(st["type"]), pr, s.n,
                        radius=dev(st["radius"]) if st["radius"] is not None else None)
    gpu_ctx.synchronize()
    for k, t in (("pos", pos), ("vel", vel), ("theta", th), ("omega", om)):
        g = host(t)
        assert rel_err(g, out_ref[k]) <= TOL, k
        assert bits_equal(g, out_ref[k]), k
    assert bits_equal(host(pr), pres_ref)


def test_reference_signatures(gpu_ctx, oracle):
    """The reference-shaped entry points (type radii, no radius arrays) on the
    triangle pile of pilha.dsml."""
    s = SCENES["pile_wrap_attract"]
    gpu_ctx.load_scene(s)
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    system, hh, ii, hs, is_, cs_ref, ce_ref, st = oracle_sorted(oracle, s)
    n, C = s.n, s.total_cells
    hash_, index = dev(np.zeros(n, np.uint32)), dev(np.zeros(n, np.uint32))
    assert lib.dem_calc_hash(h, dev(s.pos).data_ptr(), index.data_ptr(), hash_.data_ptr(), n) == 0
    assert lib.dem_sort_particles(h, hash_.data_ptr(), index.data_ptr(), n) == 0
    out = api.DeviceState(n, False)
    cs, ce = dev(np.zeros(C, np.uint32)), dev(np.zeros(C, np.uint32))
    old = {k: dev(v) for k, v in state_dict(s).items() if v is not None}
    assert lib.dem_reorder_data_and_find_cell_start(
        h, cs.data_ptr(), ce.data_ptr(), out.pos.data_ptr(), out.vel.data_ptr(), out.theta.data_ptr(),
        out.omega.data_ptr(), out.id.data_ptr(), out.loc.data_ptr(), out.type.data_ptr(), hash_.data_ptr(),
        index.data_ptr(), old["pos"].data_ptr(), old["vel"].data_ptr(), old["theta"].data_ptr(),
        old["omega"].data_ptr(), old["id"].data_ptr(), old["type"].data_ptr(), n, C) == 0
    acc, alpha, pres = dev(np.zeros((n, 2), np.float32)), dev(np.zeros(n, np.float32)), dev(np.zeros(n, np.float32))
    assert lib.dem_collide(h, out.pos.data_ptr(), out.vel.data_ptr(), acc.data_ptr(), out.omega.data_ptr(),
                           alpha.data_ptr(), out.type.data_ptr(), cs.data_ptr(), ce.data_ptr(), n, C,
                           pres.data_ptr()) == 0
    gpu_ctx.synchronize()
    assert np.array_equal(host(hash_, np.uint32), hs)
    assert np.array_equal(host(index, np.uint32), is_)
    assert np.array_equal(host(cs, np.uint32), cs_ref)
    acc_ref, alpha_ref, pres_ref = oracle.collide(system, st, cs_ref, ce_ref)
    assert rel_err(host(acc), acc_ref) <= TOL
    assert rel_err(host(pres), pres_ref) <= TOL
    assert lib.dem_integrate_system(h, out.pos.data_ptr(), out.vel.data_ptr(), acc.data_ptr(),
                                    out.theta.data_ptr(), out.omega.data_ptr(), alpha.data_ptr(),
                                    out.type.data_ptr(), pres.data_ptr(), n) == 0
    gpu_ctx.synchronize()
    ref, _ = oracle.integrate(system, st, acc_ref, alpha_ref, pres_ref)
    assert rel_err(host(out.pos), ref["pos"]) <= TOL


def test_error_codes(gpu_ctx):
    """Bad arguments come back as status codes with a message, never a crash."""
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    s = sc.column_collapse(cols=4, rows=4)
    gpu_ctx.load_scene(s)
    assert lib.dem_calc_hash(h, None, None, None, 16) == -1
    assert b"NULL" in lib.dem_last_error(h)
    x = dev(np.zeros(64, np.uint32))
    # numCells smaller than the configured grid
    assert lib.dem_collide(h, x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(),
                           x.data_ptr(), x.data_ptr(), 16, 10, x.data_ptr()) == -1
    bad = api.system_props(s)
    bad.num_cells = 7
    with pytest.raises(api.DemError):
        gpu_ctx.set_params(bad, api.particle_props(s, True))
    bad = api.system_props(s)
    bad.flags = 0  # 91 x 46 cannot wrap
    with pytest.raises(api.DemError):
        gpu_ctx.set_params(bad, api.particle_props(s, True))
    gpu_ctx.load_scene(s)


def test_step_variant_errors(gpu_ctx):
    """Exclusive candidate modes, particles a lane other than 1, 2, 4 and
    unknown flags are refused with a message; the context keeps working."""
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    for flags in (0x2 | 0x20, 0x4 | 0x20, 0x2 | 0x4, 3 << 8, 3 << 12, 0x4000):
        assert lib.dem_set_step_variant(h, flags) == -1, hex(flags)
        assert lib.dem_last_error(h)
    with pytest.raises(ValueError):
        gpu_ctx.set_step_variant(per_lane={"hash": 3})
    gpu_ctx.set_step_variant(contact_lists=True, per_lane={"hash": 4, "reorder": 2, "integrate": 4})
    gpu_ctx.set_step_variant()
