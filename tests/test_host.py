"""Host-side model code of libhzo.so (runs without a GPU).

Pins: reference decks and data files (tests/golden), scipy's independent
Euler-angle rotations, numpy brute force for grain labels and file ordering,
closed-form non-dimensional scales, glibc rand() through ctypes, and the
published SplitMix64 output sequence.
"""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import cmsl_amd as H
from cmsl_amd import _lib as L
from cmsl_amd import film as F
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(__file__))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "hzo.h")).read()
    declared = set(re.findall(r"\b(hzo_[A-Za-z0-9_]+)\s*\(", hdr))
    lib = L.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(L.SIGNATURES), declared ^ set(L.SIGNATURES)


def test_device_count_and_create_without_gpu_fail_cleanly():
    n = H.device_count()
    if n == 0:
        import ctypes as C
        cfg = L.default_config(nx=8, ny=8, nz=4)
        h = C.c_void_p()
        rc = L.lib().hzo_create(C.byref(cfg), C.byref(h))
        assert rc == -2 and not h.value
        assert b"no HIP device" in L.lib().hzo_last_error()


@pytest.mark.parametrize("kw,msg", [
    (dict(cg_fused=5), b"cg_fused"),
    (dict(cg_fused=-1), b"cg_fused"),
    (dict(tile_kz=32, cg_fused=2), b"tile_kz"),
    (dict(tile_kz=32, cg_fused=0), b"tile_kz"),
    (dict(tile_kz=16, cg_fused=0), b"tile_kz"),
    (dict(tile_kz=64, precision=L.PREC_FP64), b"tile_kz"),
    (dict(tile_kz=12), b"tile_kz"),
])
def test_create_rejects_solver_and_tile_combinations(kw, msg):
    # argument checks run before any device call, so they hold without a GPU:
    # FP32-refined depth 16 needs the single-reduction CG, depth 32 its
    # two-stencil form; FP64 takes 16 and 32 (its z-streaming kernels)
    # (cg_fused 4, the resident form, is checked against the device)
    import ctypes as C
    base = dict(nx=8, ny=8, nz=4, precision=L.PREC_FP32_REFINED)
    base.update(kw)
    cfg = L.default_config(**base)
    h = C.c_void_p()
    rc = L.lib().hzo_create(C.byref(cfg), C.byref(h))
    assert rc == -1 and not h.value
    assert msg in L.lib().hzo_last_error()


@pytest.mark.parametrize("n,grid,reps", [(0, 512, 1), ((1 << 28) + 1, 512, 1), (1 << 21, 0, 1), (1 << 21, 512, 0)])
def test_stream_probe_rejects_bad_sizes(n, grid, reps):
    # the measurement library bench.py loads (not the product library): its
    # 32-bit buffer offsets cover at most 2^28 four-byte elements; the
    # argument check runs before any device call
    import ctypes as C
    import os
    probe = C.CDLL(os.path.join(os.path.dirname(L.LIB_PATH), "libhzo_probe.so"))
    probe.hzo_probe_stream.argtypes = [C.c_longlong, C.c_int, C.c_int, C.c_void_p]
    probe.hzo_probe_last_error.restype = C.c_char_p
    us = C.c_double(-1.0)
    rc = probe.hzo_probe_stream(C.c_longlong(n), grid, reps, C.byref(us))
    assert rc == -1 and us.value == -1.0
    assert b"hzo_probe_stream" in probe.hzo_probe_last_error()


def test_product_library_holds_no_measurement_probe():
    assert not hasattr(L.lib(), "hzo_probe_stream")


def test_splitmix64_published_sequence():
    # SplitMix64 seeded with 0 (Steele, Lea, Flood 2014; Vigna's reference code)
    out = F.splitmix64(0, 3)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def _deck(path):
    d = L.Deck()
    rc = L.lib().hzo_read_deck(path.encode(), ctypes.byref(d))
    return rc, d


def test_deck_parser_on_shipped_decks():
    rc, d = _deck(os.path.join(GOLD, "InputParams_inputs"))
    assert rc == 0
    assert (d.nx, d.ny, d.nz, d.dep_flag, d.num_steps) == (940, 940, 10, 0, 800)
    assert (d.del_h, d.del_t, d.delEap, d.Eap) == (1.0, 0.01, 1.5e7, 3.0e8)
    assert (d.L_de, d.kappa_de, d.kappa_fe, d.L_fe, d.kappa_afe, d.L_afe) == (1.0, 2.5e-9, 5.066e-10, 0.833, 1e-9, 0.2)
    assert (d.dimx, d.dimy, d.dimz, d.n_layers, d.layer_counts[0]) == (8, 4, 4, 1, 10)
    assert (d.eps_fe, d.eps_afe, d.eps_de, d.G0, d.omg, d.PsNoise) == (30.0, 40.0, 15.0, 1.0, 0.075, 0.01)
    assert d.crystal_type == b"poly" and d.z_bc_type == b"diri"
    rc, d2 = _deck(os.path.join(GOLD, "InputParams_src"))
    assert rc == 0 and (d2.del_t, d2.delEap, d2.num_steps, d2.L_de, d2.L_afe, d2.eps_afe) == (0.02, 0.25e6, 50, 0.01, 0.1, 50.0)


def test_deck_layer_count_mismatch_is_an_error(tmp_path):
    p = tmp_path / "deck"
    p.write_text("nz = 12\nn_layers = 2\nlayer_counts = 5 6\n")
    rc, _ = _deck(str(p))
    assert rc == L.lib().hzo_version() * 0 - 6  # HZO_ERR_DATA
    p.write_text("nz = 11\nn_layers = 2\nlayer_counts = 5\n")
    assert _deck(str(p))[0] == -6


def test_gchar_and_rotations_match_scipy():
    ng = int(open(os.path.join(GOLD, "gnumb_poly.txt")).read().split()[0])
    assert ng == 800
    cent = np.zeros(3 * ng)
    eul = np.zeros(3 * ng)
    L.call("hzo_read_gchar", os.path.join(GOLD, "gchar_poly.txt").encode(), ng, L.ptr(cent), L.ptr(eul))
    raw = np.loadtxt(os.path.join(GOLD, "gchar_poly.txt"))
    assert np.array_equal(cent.reshape(ng, 3)[raw[:, 0].astype(int)], raw[:, 1:4])
    TR = F.euler_to_TR(eul).reshape(ng, 3, 3)
    # Bunge (phi1, Phi, phi2) passive rotation = transpose of scipy's intrinsic ZXZ
    R = Rotation.from_euler("ZXZ", eul.reshape(ng, 3)).as_matrix()
    assert np.allclose(TR, np.transpose(R, (0, 2, 1)), atol=1e-14)
    assert np.allclose(TR @ np.transpose(TR, (0, 2, 1)), np.eye(3), atol=1e-14)
    assert np.allclose(np.linalg.det(TR), 1.0, atol=1e-14)


def test_gdata_roundtrip_and_z_fastest_order(tmp_path):
    nx, ny, nz = 5, 4, 3
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    val = (100 * i + 10 * j + k).astype(np.int32)  # file order: i slowest, k fastest
    path = tmp_path / "gdata.bin"
    with open(path, "wb") as fh:
        fh.write(val.reshape(-1).tobytes())
        fh.write(np.ones(nx * ny * nz).tobytes())
    om = np.zeros(nx * ny * nz, np.int32)
    eta = np.zeros(nx * ny * nz)
    L.call("hzo_read_gdata", str(path).encode(), nx, ny, nz, L.ptr(om), L.ptr(eta))
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    assert np.array_equal(om, (100 * ii + 10 * jj + kk).reshape(-1))
    assert np.array_equal(om, O.reorder_zf_to_c(nx, ny, nz, val.reshape(-1)))
    assert np.all(eta == 1.0)
    p2 = tmp_path / "g2.bin"
    L.call("hzo_write_gdata", str(p2).encode(), nx, ny, nz, L.ptr(om))
    assert open(p2, "rb").read() == open(path, "rb").read()
    short = tmp_path / "short.bin"
    short.write_bytes(b"\0" * 10)
    assert L.lib().hzo_read_gdata(str(short).encode(), nx, ny, nz, L.ptr(om), None) == -3


def _brute_labels(nx, ny, nz, z0, span, cent):
    i, j, h = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(z0, z0 + span), indexing="ij")
    d2 = []
    for (cx, cy, cz) in cent:
        dx = cx - i
        dx = np.where(dx > 0.5 * nx, dx - nx, np.where(dx < -0.5 * nx, dx + nx, dx))
        dy = cy - j
        dy = np.where(dy > 0.5 * ny, dy - ny, np.where(dy < -0.5 * ny, dy + ny, dy))
        dz = (cz - z0) - (h - z0)
        dz = np.where(dz > 0.5 * span, dz - span, np.where(dz < -0.5 * span, dz + span, dz))
        d2.append(dx * dx + dy * dy + dz * dz)
    return np.argmin(np.stack(d2), axis=0)  # first minimum = lowest index on ties


def test_voronoi_labels_match_brute_force_and_oracle():
    nx, ny, nz, ng = 23, 17, 9, 14
    lc = [4, 5]
    om, cent, eul = F.voronoi(nx, ny, nz, ng, seed=99, layer_counts=lc)
    om3 = om.reshape(nz, ny, nx)
    # grains per layer: llround(span/nz*ng) (voronoi.c:59-66) -> 6, 8
    n0 = int(round(4 / 9 * 14))
    assert n0 == 6
    assert set(np.unique(om3[:4])) <= set(range(0, 6)) and set(np.unique(om3[4:])) <= set(range(6, 14))
    goff, z0 = 0, 0
    zf = np.zeros(nx * ny * nz, np.int32)
    for span, nL in zip(lc, [6, 8]):
        lab = _brute_labels(nx, ny, nz, z0, span, cent[goff:goff + nL]) + goff  # (nx, ny, span)
        assert np.array_equal(om3[z0:z0 + span], np.transpose(lab, (2, 1, 0)))
        O.voronoi_layer(nx, ny, nz, z0, span, np.ascontiguousarray(cent[:, 0]), np.ascontiguousarray(cent[:, 1]),
                        np.ascontiguousarray(cent[:, 2]), nL, goff, zf)
        goff += nL
        z0 += span
    assert np.array_equal(O.reorder_zf_to_c(nx, ny, nz, zf), om)
    # centroids are integers inside their layer; orientations in Bunge ranges
    assert np.all(cent == np.round(cent))
    assert np.all((cent[:6, 2] >= 0) & (cent[:6, 2] < 4)) and np.all((cent[6:, 2] >= 4) & (cent[6:, 2] < 9))
    assert np.all((eul[:, 1] >= 0) & (eul[:, 1] <= np.pi))


def test_voronoi_orientations_are_uniform_on_so3():
    _, _, eul = F.voronoi(8, 8, 4, 4000, seed=7)
    # uniform SO(3): cos(Phi) ~ U(-1,1), phi1 and phi2 ~ U(0, 2 pi)
    from scipy import stats
    assert stats.kstest(np.cos(eul[:, 1]), "uniform", args=(-1, 2)).pvalue > 1e-3
    assert stats.kstest(eul[:, 0], "uniform", args=(0, 2 * np.pi)).pvalue > 1e-3


def _py_assign(omega, centz, ng, layer_counts, fracs, seed):
    """Independent statement of the greedy assignment using glibc rand via ctypes."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    vox = np.bincount(omega[(omega >= 0) & (omega < ng)], minlength=ng)
    zs = np.concatenate([[0], np.cumsum(layer_counts)])
    mat = -np.ones(ng, int)
    for Lr in range(len(layer_counts)):
        A, Fr, D = fracs[Lr]
        s = A + Fr + D
        A, Fr, D = A / s, Fr / s, D / s
        lst = [g for g in range(ng) if zs[Lr] <= centz[g] < zs[Lr + 1]]
        tot = sum(vox[g] for g in lst)
        tA, tD = int(np.floor(A * tot + 0.5)), int(np.floor(D * tot + 0.5))
        for i in range(len(lst)):
            r = i + libc.rand() % (len(lst) - i)
            lst[i], lst[r] = lst[r], lst[i]
        p, acc = 0, 0
        while p < len(lst) and acc < tD:
            mat[lst[p]] = 2
            acc += vox[lst[p]]
            p += 1
        acc = 0
        while p < len(lst) and acc < tA:
            mat[lst[p]] = 0
            acc += vox[lst[p]]
            p += 1
        for g in lst[p:]:
            mat[g] = 1
    return mat


def test_phase_assignment_follows_greedy_rule():
    nx, ny, nz, ng = 40, 36, 10, 60
    lc = [6, 4]
    om, cent, _ = F.voronoi(nx, ny, nz, ng, seed=4, layer_counts=lc)
    fr = [(0.2, 0.6, 0.2), (0.5, 0.2, 0.3)]
    gm = F.assign_phases(nx, ny, nz, om, cent[:, 2].copy(), fr, layer_counts=lc, seed=123456)
    assert np.array_equal(gm, _py_assign(om, cent[:, 2], ng, lc, fr, 123456))
    vox = np.bincount(om, minlength=ng)
    # DE voxels reach the target with the last grain added (greedy overshoot < one grain)
    layer0 = cent[:, 2] < 6
    tot = vox[layer0].sum()
    de = vox[layer0 & (gm == 2)].sum()
    assert de >= round(0.2 * tot) and de - vox[layer0 & (gm == 2)].max() < round(0.2 * tot)


def test_phase_assignment_rejects_grain_outside_layers():
    om = np.zeros(4 * 4 * 4, np.int32)
    gm = np.zeros(1, np.int32)
    rc = L.lib().hzo_assign_phases(4, 4, 4, L.ptr(om), 1, L.ptr(np.array([9.0])), 1,
                                   L.ptr(np.array([4], np.int32)), L.ptr(np.array([0.0, 1.0, 0.0])),
                                   123456, L.ptr(gm))
    assert rc == -6


def test_interface_flag_matches_oracle():
    nx, ny, nz, ng = 20, 18, 8, 30
    om, cent, _ = F.voronoi(nx, ny, nz, ng, seed=12, layer_counts=[3, 5])
    gm = F.assign_phases(nx, ny, nz, om, cent[:, 2].copy(), (0.3, 0.4, 0.3), layer_counts=[3, 5])
    fl = F.interface_flag(nx, ny, nz, om, gm)
    c = O.cfg(nx, ny, nz, 0.5, 0.01)
    ofl = np.zeros_like(fl)
    O.lib().hzo_oracle_interface_flag(ctypes.byref(c), O._p(om), O._p(gm), O._p(ofl))
    assert np.array_equal(fl, ofl) and fl.any()
    ph = gm[om].reshape(nz, ny, nx)
    f3 = fl.reshape(nz, ny, nx)
    assert not f3[0].any() and not f3[-1].any()
    assert np.all(ph[f3 == 1] == 1)


def test_nondim_known_answers():
    # FE reference alpha=-1e9, beta=1e10, gamma=1e11, g=0 (hand-computed closed forms)
    rows, sc = F.nondim(F.BASE_RAW, L.HZO_FE)
    assert sc["P0"] == pytest.approx(0.8044958641907104, rel=1e-15)
    assert sc["E0"] == pytest.approx(2.0112396604767757e8, rel=1e-15)
    assert sc["eps0_nd"] == pytest.approx(451.77320984865605, rel=1e-15)
    inp_field = 1e8 / sc["E0"]
    assert inp_field == pytest.approx(0.4972057878785785, rel=1e-15)
    assert 0.01 / sc["P0"] == pytest.approx(0.01243014469696446, rel=1e-15)
    fe = rows[L.HZO_FE]
    assert fe["alpha"] == -1.0 and fe["a0"] == 4.0
    assert fe["kappa"] == pytest.approx(0.1, rel=1e-15)
    assert fe["L"] == pytest.approx(0.25 * 0.833 * 1e9 * 3e-9, rel=1e-15)
    de = rows[L.HZO_DE]
    assert de["beta"] == 0.0 and de["gamma"] == 0.0 and de["g"] == 0.0  # all-zero branch


def test_init_polarization_by_phase():
    nx, ny, nz = 6, 5, 4
    om = np.repeat(np.arange(3, dtype=np.int32), nx * ny * nz // 3 + 1)[: nx * ny * nz].copy()
    om[0] = -1
    gm = np.array([0, 1, 2], np.int32)
    P = F.init_polarization(nx, ny, nz, om, gm, 0.05, seed=3).reshape(6, -1)
    u = F.uniform01(3, nx * ny * nz)
    fe = om == 1
    assert np.array_equal(P[0, fe], 0.05 * (2 * u[fe] - 1)) and np.array_equal(P[0], P[2])
    assert np.all(P[3:, om == 0] == 1.0) and np.all(P[:3, om == 0] == 0.0)
    assert np.all(P[:, om == 2] == 0.0) and np.all(P[:, 0] == 0.0)


def test_sweep_schedule_matches_shipped_sweep_file():
    # sweep_sim001.txt: Eap = 3e8, delEap = 1.5e7, num_steps = 800 (inputs/InputParams:7-9).
    # Column 1 (E in MV/cm, %le) and column 12 (step count) follow from the
    # schedule alone.  Whether the level at exactly +-Eap turns the sweep
    # depends on the rounding of the accumulated field; E0 = 1.5e8 V/m sums
    # past maxEz like the reference's run did.
    ref = np.loadtxt(os.path.join(GOLD, "sweep_sim001.txt"))
    p = F.sweep_params({"E0": 1.5e8, "P0": 1.0}, 3.0e8, 1.5e7, 800)
    E, cnt = F.sweep_schedule(p)
    assert len(E) == len(ref) == 81
    assert np.array_equal(cnt, ref[:, 11].astype(int))
    # %le print precision; the two zero crossings show the accumulation residue
    # (~1e-16, its bits depend on E0)
    assert np.allclose(E, ref[:, 0], rtol=1e-6, atol=1e-12)
    # with E0 = 2e8 the accumulated 20*delEz falls short of maxEz: the turn
    # happens one level later, exactly as the reference's `>` test would do
    p2 = F.sweep_params({"E0": 2.0e8, "P0": 1.0}, 3.0e8, 1.5e7, 800)
    E2, cnt2 = F.sweep_schedule(p2)
    assert cnt2[0] == 17600 and abs(E2[0] - 3.15) < 1e-9


def test_multi_structure_layout_round_trip(tmp_path):
    nx, ny, nz = 9, 7, 4
    assert F.count_structures(tmp_path) == 0
    made = []
    for s in (1, 2, 3):
        om, cent, eul = F.voronoi(nx, ny, nz, 5, 40 + s)
        F.write_structure(tmp_path, s, nx, ny, nz, om, cent, eul)
        made.append((om, cent, eul))
    (tmp_path / "struc5").mkdir()  # a gap ends the count, as in the reference scan
    assert F.count_structures(tmp_path) == 3
    assert (tmp_path / "gnumb_poly.txt").read_text().split() == ["5"]
    for s, (om, cent, eul) in enumerate(made, start=1):
        om2, cent2, eul2 = F.load_structure(tmp_path, s, nx, ny, nz)
        assert np.array_equal(om, om2)
        # gchar holds "%lf" (6 decimals)
        assert np.allclose(cent, cent2, atol=5e-7) and np.allclose(eul, eul2, atol=5e-7)


def _boundary_mask(om, nx, ny, nz):
    g = om.reshape(nz, ny, nx)
    b = np.zeros_like(g, dtype=bool)
    for ax, per in ((2, True), (1, True)):
        for sh in (1, -1):
            b |= np.roll(g, sh, axis=ax) != g
    b[1:] |= g[1:] != g[:-1]
    b[:-1] |= g[:-1] != g[1:]
    return b.reshape(-1)


def test_grain_boundary_cells_pick_a_seeded_share_of_boundary_cells():
    nx, ny, nz, ng = 24, 20, 6, 9
    om0, _, _ = F.voronoi(nx, ny, nz, ng, 5)
    gb = _boundary_mask(om0, nx, ny, nz)
    N = om0.size

    def run(frac, seed):
        om = om0.copy()
        parent = np.zeros(ng, np.int32)
        nn = ctypes.c_int(0)
        L.call("hzo_grain_boundary_cells", nx, ny, nz, L.ptr(om), ng, frac, ctypes.c_uint64(seed), L.ptr(parent),
               ctypes.byref(nn))
        return om, parent[: nn.value - ng]

    om, par = run(0.05, 1)
    moved = om != om0
    assert moved.sum() == round(0.05 * N) and np.all(gb[moved])           # only boundary cells, the asked share
    assert np.array_equal(par[om[moved] - ng], om0[moved])                  # each new grain maps to its parent
    assert len(set(par.tolist())) == par.size and om[moved].min() == ng     # one new grain per parent
    om_again, _ = run(0.05, 1)
    assert np.array_equal(om, om_again)
    om_other, _ = run(0.05, 2)
    assert not np.array_equal(om, om_other)
    om_all, _ = run(0.99, 1)                                                # more than there are: all of them
    assert np.array_equal(om_all != om0, gb)


def test_film_inputs_with_dielectric_boundary_cells():
    inp = H.make_film_inputs(24, 20, 6, 9, fracs=(0.3, 0.5, 0.2), seed=4, gb_dielectric=0.02)
    base = H.make_film_inputs(24, 20, 6, 9, fracs=(0.3, 0.5, 0.2), seed=4)
    moved = inp.omega != base.omega
    assert moved.sum() == round(0.02 * inp.N)
    new = np.unique(inp.omega[moved])
    assert np.all(inp.grain_phase[new] == L.HZO_DE)
    par = base.omega[moved]
    assert np.array_equal(inp.TR.reshape(-1, 9)[inp.omega[moved]], base.TR.reshape(-1, 9)[par])
