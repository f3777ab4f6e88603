"""GPU parity of the FP64 drop-in path: bit-for-bit against the reference
fixtures (written by the reference's own classes) and against the oracle on
the same seeded inputs, for both step orders, walls, tau != 1 and every
kernel variant."""
import numpy as np
import pytest

import mesobasic_amd as mb
from mesobasic_amd import lbm
import oracle
from conftest import FIXTURES, load_fixture, rel_l2, same_bits

pytestmark = pytest.mark.gpu

KERNELS = [mb.KERNEL_TWO_PASS, mb.KERNEL_FUSED]


def fused_fits(q, shape):
    """Whether any fused FP64 shape takes the lattice: D2Q9 row segments
    need two x rows; D3Q19 tiles need two x planes and a y-z cross-section
    of at least one tile (8 x 32 or 4 x 64, partial or whole; whole z rows
    of 256 or 512)."""
    if q == 9:
        return shape[0] >= 2
    nx, ny, nz = shape
    return nx >= 2 and ((ny >= 8 and nz >= 32) or (ny >= 4 and nz >= 64) or nz in (256, 512))


def kernels_for(q, shape):
    return KERNELS if fused_fits(q, shape) else [mb.KERNEL_TWO_PASS]


def _lattice(meta, kernel):
    # any refusal of a lattice the kernel should take fails the test
    return mb.Lattice(meta["q"], tuple(meta["shape"]), meta["tau"], meta["G"],
                      order=meta["order"], walls=meta["walls"], precision=mb.FP64, kernel=kernel)


def _fixture_params():
    out = []
    for name in FIXTURES:
        meta, _ = load_fixture(name)
        out += [(name, k) for k in kernels_for(meta["q"], meta["shape"])]
    return out


def test_device_exp_matches_libm():
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.exp.restype = ctypes.c_double
    libm.exp.argtypes = [ctypes.c_double]
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-40, 0, 200_000), rng.uniform(-511, 511, 50_000),
                        np.array([0.0, -0.0, np.inf, -np.inf, 1e-300, 709.78, 709.79])])
    y = lbm.probe_exp(x)
    ref = np.array([libm.exp(v) for v in x])
    bad = np.flatnonzero(y.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, f"{bad.size} mismatches, first x={x[bad[0]]!r}"


@pytest.mark.parametrize("name,kernel", _fixture_params())
def test_fixture_bitwise(name, kernel):
    meta, data = load_fixture(name)
    print(f"case {name} kernel {kernel}", flush=True)
    lat = _lattice(meta, kernel)
    lat.set_rho(data["rho0"])
    lat.init()
    cur = 0
    for s in sorted(meta["steps"]):
        lat.step(s - cur)
        cur = s
        got = lat.fields(mb.FIELDS_REFERENCE)
        for k in ("rho", "u", "v") + (("w",) if meta["q"] == 19 else ()):
            assert same_bits(got[k].ravel(), data[f"{k}_{s}"]), (name, s, k,
                                                               rel_l2(got[k], data[f"{k}_{s}"]))
        if s in meta["steps_f"]:
            assert same_bits(lat.get_f().ravel(), data[f"f_{s}"]), (name, s, "f")
    lat.close()


CASES = [
    # q, shape, tau, G, order, walls, steps
    (19, (9, 7, 6), 1.0, -5.5, 2, False, 25),
    (19, (9, 7, 6), 0.65, -5.0, 3, True, 25),
    (19, (6, 10, 7), 1.3, -5.5, 2, True, 25),
    (9, (17, 12), 1.0, -5.0, 3, False, 40),
    (9, (11, 16), 0.7, -4.5, 3, True, 40),
    (9, (33, 31), 1.0, -5.0, 2, False, 200),
]


# lattices the fused kernel takes: tile multiples (nx = 70 spans two full
# 32-plane chunks and a partial one), then partial tiles
FUSED_CASES = [
    (19, (6, 16, 32), 1.0, -5.5, 3, False, 30),
    (19, (5, 32, 32), 0.7, -5.0, 3, True, 30),
    (19, (70, 16, 32), 1.0, -5.5, 3, False, 12),
    (19, (3, 16, 64), 1.2, -5.5, 2, True, 20),
    (19, (5, 6, 256), 1.0, -5.5, 3, False, 15),
    (19, (3, 2, 512), 0.9, -5.0, 3, True, 12),
    # partial tiles (ny, nz not multiples of the tile)
    (19, (7, 12, 90), 1.0, -5.5, 3, False, 20),
    (19, (6, 21, 70), 0.8, -5.0, 2, True, 15),
]


@pytest.mark.parametrize("case,kernel", [(c, k) for c in CASES + FUSED_CASES
                                         for k in kernels_for(c[0], c[1])])
def test_matches_oracle_bitwise(case, kernel):
    q, shape, tau, G, order, walls, steps = case
    print(f"case {case} kernel {kernel}", flush=True)
    rho0 = mb.noise_rho(shape, 21)
    o = oracle.Oracle(q, shape, tau, G, order=order, walls=walls)
    o.set_rho(rho0)
    o.init()
    meta = dict(q=q, shape=shape, tau=tau, G=G, order=order, walls=walls)
    lat = _lattice(meta, kernel)
    lat.set_rho(rho0)
    lat.init()
    for chunk in (1, 1, steps - 2):
        o.step(chunk)
        lat.step(chunk)
        ref = o.fields()
        got = lat.fields(mb.FIELDS_REFERENCE)
        for k in ("rho", "u", "v", "w"):
            assert same_bits(got[k], ref[k]), (k, rel_l2(got[k], ref[k]))
        assert same_bits(lat.get_f(), ref["f"])
        cur = lat.fields(mb.FIELDS_CURRENT)
        mom = o.moments()
        for k in ("rho", "u", "v", "w"):
            assert same_bits(cur[k], mom[k]), k
    lat.close()


@pytest.mark.parametrize("shape", [(40, 48, 64), (128, 32, 128)])
def test_fused_equals_two_pass_bitwise(shape):
    # (128, 32, 128): 2 x 4 tiles x 4 chunks, so the XCD-ordered tile map is used
    rho0 = mb.droplet_rho(shape, 10.0)
    a = mb.Lattice(19, shape, 1.0, -5.5, kernel=mb.KERNEL_TWO_PASS)
    b = mb.Lattice(19, shape, 1.0, -5.5, kernel=mb.KERNEL_FUSED)
    assert b.kernel == mb.KERNEL_FUSED
    for lat in (a, b):
        lat.set_rho(rho0); lat.init(); lat.step(25)
    assert same_bits(a.get_f(), b.get_f())
    fa, fb = a.fields(), b.fields()
    for k in ("rho", "u", "v", "w"):
        assert same_bits(fa[k], fb[k]), k


@pytest.mark.parametrize("shape", [(24, 32, 64), (6, 4, 256), (4, 3, 512)])
def test_fused_every_tile_shape_and_chunk_bitwise(shape):
    # results must not depend on the launch shape; (6, 4, 256) and (4, 3, 512)
    # take the full-row tiles, (24, 32, 64) the z-tiled ones
    rho0 = mb.noise_rho(shape, 17)
    ref = mb.Lattice(19, shape, 0.8, -5.5, kernel=mb.KERNEL_TWO_PASS, walls=True)
    ref.set_rho(rho0); ref.init(); ref.step(9)
    want = ref.get_f()
    tried = 0
    for tile in range(6):
        for chunk, order, nts in ((1, 0, 0), (2, 1, 1), (5, 2, 1), (64, 2, 0)):
            lat = mb.Lattice(19, shape, 0.8, -5.5, kernel=mb.KERNEL_FUSED, walls=True)
            try:
                lat.set_tuning(tile, chunk, order, nts)
            except mb.LbmError:
                lat.close()
                break           # this tile shape does not fit the lattice
            print(f"tile {tile} chunk {chunk} order {order} nts {nts}", flush=True)
            assert lat.tuning == (tile, chunk, order, nts)
            lat.set_rho(rho0); lat.init(); lat.step(9)
            assert same_bits(lat.get_f(), want), (tile, chunk)
            lat.close()
            tried += 1
    assert tried >= 4


@pytest.mark.parametrize("shape", [(9, 20, 100), (5, 13, 70), (6, 9, 65), (4, 8, 64)])
def test_fused_partial_tiles_bitwise(shape):
    # lattices whose ny or nz is not a multiple of the tile: the last tile of
    # a column wraps periodically and stores only its real nodes; (4, 8, 64)
    # is a control that every tile divides
    rho0 = mb.noise_rho(shape, 19)
    ref = mb.Lattice(19, shape, 0.8, -5.5, kernel=mb.KERNEL_TWO_PASS, walls=True)
    ref.set_rho(rho0); ref.init(); ref.step(9)
    want = ref.get_f()
    ref.close()
    tried = 0
    for tile in (3, 4, 5):
        for chunk, order, nts in ((1, 0, 0), (2, 3, 1), (5, 1, 1), (64, 3, 0)):
            lat = mb.Lattice(19, shape, 0.8, -5.5, kernel=mb.KERNEL_FUSED, walls=True)
            try:
                lat.set_tuning(tile, chunk, order, nts)
            except mb.LbmError:
                lat.close()
                break           # lattice narrower than this tile
            print(f"shape {shape} tile {tile} chunk {chunk} order {order} nts {nts}", flush=True)
            lat.set_rho(rho0); lat.init(); lat.step(9)
            assert same_bits(lat.get_f(), want), (tile, chunk, order)
            lat.close()
            tried += 1
    assert tried >= 4


ROW_SHAPES = (8, 9, 10)     # D2Q9 row segments of 256, 128 and 64 nodes


@pytest.mark.parametrize("shape,order", [((40, 300), 3), ((9, 70), 2), ((6, 5), 3),
                                         ((2, 129), 3), ((31, 256), 2), ((12, 64), 3),
                                         ((7, 128), 2)])
def test_fused_rows9_every_segment_and_chunk_bitwise(shape, order):
    # D2Q9 row march vs the two-pass kernels: segments longer than the row
    # (wrapped lanes), ragged last segments (300 = 256 + 44), full rows
    # (ny = 64 / 128 / 256 equal to a segment: no ring), nx = 2, walls on
    # rows 1 and ny-2, tau != 1; order 2 runs the LBfluid2D first step
    # two-pass and the rest fused
    rho0 = mb.noise_rho(shape, 23)
    walls = shape[1] >= 3
    ref = mb.Lattice(9, shape, 0.8, -5.0, order=order, kernel=mb.KERNEL_TWO_PASS, walls=walls)
    ref.set_rho(rho0); ref.init(); ref.step(11)
    want = ref.get_f()
    ref.close()
    for tile in ROW_SHAPES:
        for chunk, nts in ((1, 0), (2, 1), (3, 1), (7, 0), (64, 1)):
            print(f"shape {shape} tile {tile} chunk {chunk} nts {nts}", flush=True)
            lat = mb.Lattice(9, shape, 0.8, -5.0, order=order, kernel=mb.KERNEL_FUSED, walls=walls)
            lat.set_tuning(tile, chunk, 3, nts)
            assert lat.tuning == (tile, chunk, 3, nts)
            lat.set_rho(rho0); lat.init(); lat.step(11)
            assert same_bits(lat.get_f(), want), (tile, chunk)
            lat.close()


BLOCK_SHAPES = (11, 12, 13)  # D2Q9 temporal blocks: 16x16 tiles x 2 steps, 16x16 x 4, 8x8 x 4


@pytest.mark.parametrize("shape,order,tau", [((32, 48), 2, 1.0), ((64, 16), 3, 0.8), ((16, 16), 2, 1.3),
                                             ((80, 96), 2, 1.0), ((8, 24), 3, 1.0)])
def test_temporal_blocks_bitwise(shape, order, tau):
    # K steps per launch against single steps of the two-pass kernels: every
    # block shape whose tile divides the lattice (16 x 16 and 8 x 24 are
    # narrower than the 24- and 32-node regions, which then hold some nodes
    # twice), step counts that leave remainders, both step orders and tau;
    # populations, reference arrays and force after each call
    rho0 = mb.noise_rho(shape, 29)
    ref = mb.Lattice(9, shape, tau, -5.0, order=order, kernel=mb.KERNEL_TWO_PASS)
    ref.set_rho(rho0); ref.init()
    calls = (1, 6, 11, 5, 2)
    want = []
    for n in calls:
        ref.step(n)
        want.append((ref.get_f(), ref.fields(mb.FIELDS_REFERENCE), ref.force()))
    ref.close()
    tried = 0
    for block in BLOCK_SHAPES:
        t = 16 if block < 13 else 8
        if shape[0] % t or shape[1] % t:
            continue
        print(f"shape {shape} order {order} tau {tau} block {block}", flush=True)
        lat = mb.Lattice(9, shape, tau, -5.0, order=order, kernel=mb.KERNEL_FUSED)
        lat.set_tuning(block)
        assert lat.tuning[0] == block
        lat.set_rho(rho0); lat.init()
        for n, (f, fl, F) in zip(calls, want):
            lat.step(n)
            assert same_bits(lat.get_f(), f), (block, n)
            got = lat.fields(mb.FIELDS_REFERENCE)
            for k in ("rho", "u", "v"):
                assert same_bits(got[k], fl[k]), (block, n, k)
            assert all(same_bits(a, b) for a, b in zip(lat.force(), F)), (block, n)
        lat.close()
        tried += 1
    assert tried >= 1


def test_temporal_blocks_refuse_what_they_do_not_fit():
    lat = mb.Lattice(9, (40, 32), 1.0, -5.0, kernel=mb.KERNEL_FUSED)
    with pytest.raises(mb.LbmError):
        lat.set_tuning(11)                 # 40 is not a multiple of 16
    lat.set_tuning(13)                     # 8 x 8 tiles fit
    lat.close()
    walls = mb.Lattice(9, (32, 32), 1.0, -5.0, kernel=mb.KERNEL_FUSED, walls=True)
    with pytest.raises(mb.LbmError):
        walls.set_tuning(12)               # periodic lattices only
    walls.close()


def test_fused_rows9_tau1_long_run_matches_oracle():
    # tau = 1 instantiation over many steps, default launch configuration
    shape = (48, 80)
    rho0 = mb.noise_rho(shape, 5)
    o = oracle.Oracle(9, shape, 1.0, -5.0, order=2)
    o.set_rho(rho0); o.init(); o.step(150)
    lat = mb.Lattice(9, shape, 1.0, -5.0, order=2, kernel=mb.KERNEL_FUSED)
    assert lat.kernel == mb.KERNEL_FUSED
    lat.set_rho(rho0); lat.init(); lat.step(150)
    assert same_bits(lat.get_f(), o.fields()["f"])
    lat.close()


@pytest.mark.parametrize("q,shape", [(9, (1, 64)), (19, (1, 8, 64))])
def test_fused_refuses_a_single_x_plane(q, shape):
    # the march wraps x+2 with one subtraction of nx, so nx = 1 stays two-pass
    with pytest.raises(mb.LbmError):
        mb.Lattice(q, shape, 1.0, -5.0, kernel=mb.KERNEL_FUSED)
    lat = mb.Lattice(q, shape, 1.0, -5.0)
    assert lat.kernel == mb.KERNEL_TWO_PASS
    lat.close()


def test_set_f_roundtrip_and_restart():
    shape = (8, 6, 5)
    rho0 = mb.noise_rho(shape, 4)
    a = mb.Lattice(19, shape, 1.0, -5.5, kernel=mb.KERNEL_TWO_PASS)
    a.set_rho(rho0); a.init(); a.step(7)
    f7 = a.get_f()
    b = mb.Lattice(19, shape, 1.0, -5.5, kernel=mb.KERNEL_TWO_PASS)
    b.set_f(f7)
    assert same_bits(b.get_f(), f7)
    a.step(5); b.step(5)
    assert same_bits(a.get_f(), b.get_f())


@pytest.mark.parametrize("q,shape", [(19, (16, 12, 10)), (9, (64, 48))])
def test_conservation_fp64(q, shape):
    rho0 = mb.noise_rho(shape, 6)
    lat = mb.Lattice(q, shape, 1.0, -5.5 if q == 19 else -5.0)
    lat.set_rho(rho0); lat.init()
    c0 = lat.conservation()
    lat.step(100)
    c1 = lat.conservation()
    assert abs(c1[0] - c0[0]) / c0[0] < 1e-12
    # total momentum: sum of forces is zero, so momentum stays ~ round-off
    assert np.all(np.abs(c1[1:]) < 1e-9 * c0[0])
    m = lat.fields(mb.FIELDS_CURRENT)["rho"].sum()
    assert abs(m - c1[0]) / c1[0] < 1e-12


def test_errors_are_reported():
    with pytest.raises(mb.LbmError):
        mb.Lattice(19, (4, 4, 4), -1.0, -5.5)
    lat = mb.Lattice(19, (4, 4, 4), 1.0, -5.5)
    with pytest.raises(mb.LbmError):
        lat.step(1)        # before init
