"""Host-side mirror of the reference's per-simulation setup and step loop.

``Film`` owns one device context (one film) behind the C ABI of libhzo.so.
``make_film_inputs`` builds the seeded inputs of a film the way the reference
does in SetupAndRunSimulation (HZO_Code_github/src/main.cu:261-359):
non-dimensionalisation, grain structure, phase assignment (Init_Conf,
init_conf.cu:108-285), interface flag (main.cu:13-52) and initial
polarisation (init_conf.cu:14-66).  All compute runs in libhzo.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

# Reference base coefficients (SI).  FE: the single-film case (alpha, beta,
# gamma of the 64x64x16 film); AFE and DE: HZO_Code_github/src/inputs/
# SingleL_input.txt:4.  L, kappa, eps_r: HZO_Code_github/inputs/InputParams:14-36,
# with kappa_FE = 1e-10 (the gradient coefficient of the single-film case).
BASE_RAW = {
    L.HZO_AFE: dict(alpha=5.022532e8, beta=-9.065448e10, gamma=2.189509e12, g=5.45348e8, a0=8.0e8,
                    L=0.2, kappa=1.0e-9, eps_r=40.0),
    L.HZO_FE: dict(alpha=-1.0e9, beta=1.0e10, gamma=1.0e11, g=0.0, a0=1.0e9,
                   L=0.833, kappa=1.0e-10, eps_r=30.0),
    L.HZO_DE: dict(alpha=9.5e9, beta=0.0, gamma=0.0, g=0.0, a0=2.372e9,
                   L=1.0, kappa=2.5e-9, eps_r=15.0),
}


def read_landau_table(path, max_rows=100000):
    """Rows of a Landau-coefficient table (hzo_read_landau_table): 13-token
    pool rows (LoadLandauPool, main.cu:112-151), the 12-column
    FreeEnergyParams.txt and the 16-token SingleL_input.txt.  Returns
    (rows, skipped): each row a dict of SI coefficients with 'ref_type'
    (-1 without a label), 'fracs' (percent, None without) and 'ncols'."""
    n, sk = C.c_int(0), C.c_int(0)
    L.call("hzo_read_landau_table", str(path).encode(), 0, None, C.byref(n), C.byref(sk))
    buf = (L.LandauRow * max(n.value, 1))()
    L.call("hzo_read_landau_table", str(path).encode(), n.value, buf, C.byref(n), C.byref(sk))
    rows = []
    for r in buf[: n.value]:
        rows.append(dict(afe=list(r.afe), fe=list(r.fe), de=list(r.de),
                         fracs=None if r.ncols != 16 else list(r.fracs),
                         ref_type=int(r.ref_type), ncols=int(r.ncols), _c=r))
    return rows, sk.value


def landau_raw(row, base=None):
    """PickLandauSet (main.cu:155-178): the row's alpha, beta, gamma, g, a0 for
    AFE and FE and (alpha, a0) for DE over base's L, kappa, eps_r."""
    base = base or BASE_RAW
    arr = (L.PhaseRaw * 3)()
    for q in range(3):
        for k in ("alpha", "beta", "gamma", "g", "a0", "L", "kappa", "eps_r"):
            setattr(arr[q], k, float(base[q][k]))
        arr[q].transition_type = int(base[q].get("transition_type", 0))
    L.call("hzo_landau_to_raw", C.byref(row["_c"]), arr)
    return {q: {k: getattr(arr[q], k) for k in ("alpha", "beta", "gamma", "g", "a0", "L", "kappa", "eps_r")}
            for q in range(3)}


def splitmix64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    L.call("hzo_splitmix64", C.c_uint64(seed), C.c_int64(n), L.ptr(out))
    return out


def uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def nondim(raw: dict, ref_type: int = L.HZO_FE):
    """make_nd_wrt_reference for the three phases; returns (rows, scales)."""
    arr = (L.PhaseRaw * 3)()
    for q in range(3):
        r = raw[q]
        for k in ("alpha", "beta", "gamma", "g", "a0", "L", "kappa", "eps_r"):
            setattr(arr[q], k, float(r[k]))
        arr[q].transition_type = int(r.get("transition_type", 0))
    out = (L.PhaseND * 3)()
    sc = np.zeros(4)
    L.call("hzo_nondim", arr, ref_type, out, L.ptr(sc))
    rows = [{k: getattr(out[q], k) for k in ("alpha", "beta", "gamma", "g", "a0", "kappa", "L", "eps_r")}
            for q in range(3)]
    return rows, {"P0": sc[0], "E0": sc[1], "eps0_nd": sc[2], "alpha_ref": sc[3]}


def voronoi(nx, ny, nz, ng, seed, layer_counts=None, device=None):
    """Layered periodic Voronoi structure; device = a GPU index labels the
    voxels there (the same labels as the CPU labeller)."""
    layer_counts = list(layer_counts or [nz])
    lc = np.array(layer_counts, np.int32)
    omega = np.zeros(nx * ny * nz, np.int32)
    cent = np.zeros(3 * ng)
    eul = np.zeros(3 * ng)
    if device is None:
        L.call("hzo_voronoi", nx, ny, nz, len(layer_counts), L.ptr(lc), ng, C.c_uint64(seed),
               L.ptr(omega), L.ptr(cent), L.ptr(eul))
    else:
        L.call("hzo_voronoi_gpu", int(device), nx, ny, nz, len(layer_counts), L.ptr(lc), ng, C.c_uint64(seed),
               L.ptr(omega), L.ptr(cent), L.ptr(eul))
    return omega, cent.reshape(ng, 3), eul.reshape(ng, 3)


def euler_to_TR(eul: np.ndarray) -> np.ndarray:
    eul = np.ascontiguousarray(eul, np.float64).reshape(-1)
    ng = eul.size // 3
    TR = np.zeros(9 * ng)
    L.call("hzo_euler_to_TR", ng, L.ptr(eul), L.ptr(TR))
    return TR


def assign_phases(nx, ny, nz, omega, centz, fracs_afe_fe_de, layer_counts=None, seed=123456):
    layer_counts = list(layer_counts or [nz])
    lc = np.array(layer_counts, np.int32)
    fr = np.ascontiguousarray(np.broadcast_to(np.asarray(fracs_afe_fe_de, np.float64),
                                              (len(layer_counts), 3)))
    ng = int(np.asarray(centz).size)
    gm = np.zeros(ng, np.int32)
    L.call("hzo_assign_phases", nx, ny, nz, L.ptr(np.ascontiguousarray(omega, np.int32)), ng,
           L.ptr(np.ascontiguousarray(centz, np.float64)), len(layer_counts), L.ptr(lc), L.ptr(fr),
           C.c_uint(seed), L.ptr(gm))
    return gm


def interface_flag(nx, ny, nz, omega, grain_phase):
    out = np.zeros(nx * ny * nz, np.int32)
    L.call("hzo_interface_flag", nx, ny, nz, L.ptr(omega), int(grain_phase.size), L.ptr(grain_phase),
           L.ptr(out))
    return out


def init_polarization(nx, ny, nz, omega, grain_phase, amp_fe, seed):
    P = np.zeros(6 * nx * ny * nz)
    L.call("hzo_init_polarization", nx, ny, nz, L.ptr(omega), int(grain_phase.size), L.ptr(grain_phase),
           float(amp_fe), C.c_uint64(seed), L.ptr(P))
    return P


@dataclass
class FilmInputs:
    nx: int
    ny: int
    nz: int
    del_h: float
    del_t: float
    omega: np.ndarray
    grain_phase: np.ndarray
    TR: np.ndarray
    phases: list
    scales: dict
    P0: np.ndarray
    iface: np.ndarray
    centroids: np.ndarray = None
    euler: np.ndarray = None
    meta: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.nx * self.ny * self.nz

    @property
    def ng(self):
        return int(self.grain_phase.size)

    def field_nd(self, E_MV_per_cm):
        """Applied field in MV/cm -> non-dimensional (maxEz = Eap/E0_ref, main.cu:347)."""
        return np.asarray(E_MV_per_cm, np.float64) * 1e8 / self.scales["E0"]


def make_film_inputs(nx, ny, nz, ng, fracs=(0.0, 1.0, 0.0), seed=1, del_h=0.5, del_t=0.01,
                     jitter=0.0, amp_fe_SI=0.01, raw=None, ref_type=L.HZO_FE, layer_counts=None,
                     phase_seed=123456, label_device=None, gb_dielectric=0.0, landau=None) -> FilmInputs:
    """Seeded film: Voronoi grains, phase assignment, Landau jitter, initial P.

    jitter: alpha, beta, gamma of each phase are multiplied by U(1-jitter, 1+jitter)
    (one draw per phase and coefficient).  amp_fe_SI: FE initial |P| bound in C/m^2.
    label_device: a GPU index labels the Voronoi voxels there (same labels).
    gb_dielectric: share of all cells, taken from grain-boundary cells, that
    carry the dielectric phase (new grains with their parents' orientation).
    landau: a row of read_landau_table; its coefficients replace those of raw
    (PickLandauSet) and its label, when it has one, sets ref_type.
    """
    raw = {q: dict(v) for q, v in (raw or BASE_RAW).items()}
    if landau is not None:
        raw = landau_raw(landau, raw)
        if landau["ref_type"] >= 0:
            ref_type = landau["ref_type"]
    if jitter:
        u = uniform01(seed ^ 0x5EED5EED, 9).reshape(3, 3)
        for q in range(3):
            for t, k in enumerate(("alpha", "beta", "gamma")):
                raw[q][k] = raw[q][k] * (1.0 - jitter + 2.0 * jitter * u[q, t])
    rows, scales = nondim(raw, ref_type)
    omega, cent, eul = voronoi(nx, ny, nz, ng, seed, layer_counts, device=label_device)
    gphase = assign_phases(nx, ny, nz, omega, cent[:, 2].copy(), fracs, layer_counts, phase_seed)
    if gb_dielectric > 0.0:
        parent = np.zeros(ng, np.int32)
        ng_new = C.c_int(0)
        L.call("hzo_grain_boundary_cells", nx, ny, nz, L.ptr(omega), ng, float(gb_dielectric),
               C.c_uint64(seed ^ 0x6B6B6B6B), L.ptr(parent), C.byref(ng_new))
        par = parent[: ng_new.value - ng]
        gphase = np.concatenate([gphase, np.full(par.size, L.HZO_DE, np.int32)])
        cent = np.concatenate([cent, cent[par]])
        eul = np.concatenate([eul, eul[par]])
    TR = euler_to_TR(eul)
    iface = interface_flag(nx, ny, nz, omega, gphase)
    P0 = init_polarization(nx, ny, nz, omega, gphase, amp_fe_SI / scales["P0"], seed + 7)
    return FilmInputs(nx, ny, nz, del_h, del_t, omega, gphase, TR, rows, scales, P0, iface, cent, eul,
                      meta=dict(seed=seed, fracs=tuple(fracs), jitter=jitter, raw=raw))


def batch_film_specs(n, seed=2024):
    """Film shapes of the batched workload: (nx = ny from {128, 192, 256, 384},
    nz from {16, 24, 32, 48}, 50-300 grains, film seed) for n films."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice([128, 192, 256, 384])), int(rng.choice([16, 24, 32, 48])),
             int(rng.integers(50, 301)), 100 + f) for f in range(n)]


def triangle_wave(nsteps_period: int, amp: float, nsteps: int, start: int = 0) -> np.ndarray:
    """E(n) = amp*tri(n/T), tri(s) = 4s | 2-4s | 4s-4 on [0,1/4] | (1/4,3/4] | (3/4,1]."""
    s = ((np.arange(start, start + nsteps) % nsteps_period) / nsteps_period).astype(np.float64)
    tri = np.where(s <= 0.25, 4 * s, np.where(s <= 0.75, 2 - 4 * s, 4 * s - 4))
    return amp * tri


def sweep_params(scales, Eap_SI, delEap_SI, num_steps, max_levels=0) -> L.SweepParams:
    E0 = scales["E0"]
    return L.SweepParams(Eap_SI / E0, delEap_SI / E0, int(num_steps), int(max_levels), E0 / 1.0e8,
                         scales["P0"] * 100.0)


def sweep_schedule(p: L.SweepParams, max_rows=100000):
    E = np.zeros(max_rows)
    cnt = np.zeros(max_rows, np.int32)
    n = C.c_int(0)
    L.call("hzo_sweep_schedule", C.byref(p), max_rows, L.ptr(E), L.ptr(cnt), C.byref(n))
    return E[: n.value], cnt[: n.value]


def write_vector_field(path, vx, vy, vz, scale=1.0):
    """write_vector_field_binary (generate.cu:11-30)."""
    vx, vy, vz = (np.ascontiguousarray(v, np.float64) for v in (vx, vy, vz))
    L.call("hzo_write_vector_field", str(path).encode(), vx.size, L.ptr(vx), L.ptr(vy), L.ptr(vz), float(scale))


def write_grain_map(path, nx, ny, nz, grain_map, scale=1.0):
    """write_field (generate.cu:33-65): text grain map."""
    g = np.ascontiguousarray(grain_map, np.int32)
    L.call("hzo_write_grain_map", str(path).encode(), nx, ny, nz, L.ptr(g), float(scale))


def write_midplane_slice(path, nx, ny, nz, v):
    """write_midplane_slice (generate.cu:67-88)."""
    v = np.ascontiguousarray(v, np.float64)
    L.call("hzo_write_midplane_slice", str(path).encode(), nx, ny, nz, L.ptr(v))


def write_checkpoint(path, nx, ny, nz, P, E, peloop=0, count=0, Ez_ext=0.0, incEz=0.0):
    """Reference checkpoint file (generate.cu:90-168); P[6N], E[3N] SoA."""
    P = np.ascontiguousarray(P, np.float64)
    E = np.ascontiguousarray(E, np.float64)
    L.call("hzo_write_checkpoint", str(path).encode(), nx, ny, nz, int(peloop), int(count), float(Ez_ext),
           float(incEz), L.ptr(P), L.ptr(E))


def read_checkpoint(path, nx, ny, nz):
    """Reads a reference checkpoint (generate.cu:170-264): header dict, P[6N], E[3N]."""
    N = nx * ny * nz
    P = np.zeros(6 * N)
    E = np.zeros(3 * N)
    pe, cnt = C.c_int(0), C.c_int(0)
    ez, inc = C.c_double(0.0), C.c_double(0.0)
    L.call("hzo_read_checkpoint", str(path).encode(), nx, ny, nz, C.byref(pe), C.byref(cnt), C.byref(ez),
           C.byref(inc), L.ptr(P), L.ptr(E))
    return {"peloop": pe.value, "count": cnt.value, "Ez_ext": ez.value, "incEz": inc.value}, P, E


def write_raw_f64(path, v):
    """Raw f64 dump in voxel order (used for the potential)."""
    v = np.ascontiguousarray(v, np.float64)
    L.call("hzo_write_raw_f64", str(path).encode(), v.size, L.ptr(v))


def count_structures(inputs_dir) -> int:
    """Number of consecutive inputs_dir/strucN directories (main.cu:430-440)."""
    n = C.c_int(0)
    L.call("hzo_count_structures", str(inputs_dir).encode(), C.byref(n))
    return n.value


def write_structure(inputs_dir, s, nx, ny, nz, omega, cent, euler):
    """Writes structure s of the multi-structure layout: inputs_dir/strucS/
    gdata_poly.bin and gchar_poly.txt, and inputs_dir/gnumb_poly.txt."""
    import os
    d = os.path.join(str(inputs_dir), f"struc{s}")
    os.makedirs(d, exist_ok=True)
    omega = np.ascontiguousarray(omega, np.int32)
    cent = np.ascontiguousarray(cent, np.float64).reshape(-1)
    euler = np.ascontiguousarray(euler, np.float64).reshape(-1)
    ng = cent.size // 3
    L.call("hzo_write_gdata", os.path.join(d, "gdata_poly.bin").encode(), nx, ny, nz, L.ptr(omega))
    L.call("hzo_write_gchar", os.path.join(d, "gchar_poly.txt").encode(), ng, L.ptr(cent), L.ptr(euler))
    with open(os.path.join(str(inputs_dir), "gnumb_poly.txt"), "w") as fh:
        fh.write(f"{ng}\n")


def load_structure(inputs_dir, s, nx, ny, nz):
    """Structure s of the multi-structure layout (grain_io.cu:10-99): omega
    (C order), centroids (ng, 3), Euler angles (ng, 3)."""
    import os
    ng = C.c_int(0)
    L.call("hzo_read_grain_count", os.path.join(str(inputs_dir), "gnumb_poly.txt").encode(), C.byref(ng))
    d = os.path.join(str(inputs_dir), f"struc{s}")
    omega = np.zeros(nx * ny * nz, np.int32)
    L.call("hzo_read_gdata", os.path.join(d, "gdata_poly.bin").encode(), nx, ny, nz, L.ptr(omega), None)
    cent = np.zeros(3 * ng.value)
    eul = np.zeros(3 * ng.value)
    L.call("hzo_read_gchar", os.path.join(d, "gchar_poly.txt").encode(), ng.value, L.ptr(cent), L.ptr(eul))
    return omega, cent.reshape(-1, 3), eul.reshape(-1, 3)


class Film:
    """One film on one GPU: the device context of libhzo.so.

    FP32-refined solver (``cg_fused``): 0 the two-kernel PCG, 1 the
    single-reduction CG in the form chosen by film size (default), 2 that CG
    carrying w = A u and s = A p, 3 that CG forming them by stencils.
    ``tile_kz`` 0 picks the tile depth (2, 4, 8; 16 and 32 for large films).
    """

    def __init__(self, inp: FilmInputs, precision=L.PREC_FP64, pcg_rule=L.PCG_RULE_REFERENCE,
                 tol_rel=1e-2, tol_abs=1e-3, max_iter=4000, dep_flag=1, zbc=L.ZBC_DIRICHLET,
                 refine_every=20, phi=None, set_state=True, tile_kz=0, cg_fused=1, g_bot=0.0, g_top=0.0):
        self.inp = inp
        cfg = L.default_config(nx=inp.nx, ny=inp.ny, nz=inp.nz, del_h=inp.del_h, del_t=inp.del_t,
                               dep_flag=dep_flag, zbc=zbc, eps0_nd=inp.scales["eps0_nd"], g_bot=g_bot, g_top=g_top,
                               precision=precision, pcg_rule=pcg_rule, tol_abs=tol_abs, tol_rel=tol_rel,
                               max_iter=max_iter, refine_every=refine_every, tile_kz=tile_kz,
                               cg_fused=cg_fused)
        self.cfg = cfg
        h = C.c_void_p()
        L.call("hzo_create", C.byref(cfg), C.byref(h))
        self.h = h
        L.call("hzo_set_microstructure", h, L.ptr(inp.omega), inp.ng, L.ptr(inp.grain_phase), L.ptr(inp.TR),
               L.ptr(inp.iface))
        L.call("hzo_set_phase_table", h, L.phases_nd(inp.phases))
        if set_state:
            self.set_state(inp.P0, phi)

    @property
    def tile_depth(self) -> int:
        return L.lib().hzo_tile_depth(self.h)

    @property
    def cg_form(self) -> int:
        """0 two-kernel PCG, 1 single-reduction CG carrying w and s, 2 the
        single-reduction CG forming them by stencils, 3 the register-resident
        single-reduction CG (one launch per solve)."""
        return L.lib().hzo_cg_form(self.h)

    def close(self):
        if getattr(self, "h", None):
            L.lib().hzo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_phase_table(self, rows):
        """Per-phase non-dimensional coefficients (rows of dicts or 8-tuples, AFE/FE/DE)."""
        L.call("hzo_set_phase_table", self.h, L.phases_nd(rows))

    # ---------------------------------------------------------- reverse mode
    def adjoint_reserve(self, nsteps: int, ckpt_every: int = 50, tol: float = 1e-8):
        """Device buffers for a recorded run of up to nsteps steps."""
        L.call("hzo_adjoint_reserve", self.h, int(nsteps), int(ckpt_every), float(tol))
        self._adj_steps = int(nsteps)

    def forward_record(self, Ez) -> np.ndarray:
        """Runs len(Ez) steps from the current state, checkpointing; returns
        the loop samples m_n = <Pmz> after each step."""
        Ez = np.ascontiguousarray(Ez, np.float64)
        m = np.zeros(Ez.size)
        L.call("hzo_forward_record", self.h, L.ptr(Ez), int(Ez.size), L.ptr(m))
        return m

    def backward(self, dL_dm):
        """Gradient of L with respect to (alpha, beta, gamma, kappa) of each
        phase, shape (3, 4) [AFE, FE, DE], for the last recorded run; also
        returns the adjoint PCG iteration count."""
        dL_dm = np.ascontiguousarray(dL_dm, np.float64)
        g = np.zeros(12)
        its = C.c_longlong(0)
        L.call("hzo_backward", self.h, L.ptr(dL_dm), L.ptr(g), C.byref(its))
        return g.reshape(3, 4), int(its.value)

    def set_state(self, P, phi=None):
        P = np.ascontiguousarray(P, np.float64)
        L.call("hzo_set_state", self.h, L.ptr(P), L.ptr(None if phi is None else np.ascontiguousarray(phi)))

    def get_state(self):
        N = self.inp.N
        P = np.empty(6 * N)
        phi = np.empty(N)
        E = np.empty(3 * N)
        L.call("hzo_get_state", self.h, L.ptr(P), L.ptr(phi), L.ptr(E))
        return P, phi, E

    def tdgl_step(self, Ez_ext: float):
        L.call("hzo_tdgl_step", self.h, float(Ez_ext))

    def depol_solve(self) -> L.PcgStats:
        st = L.PcgStats()
        L.call("hzo_depol_solve", self.h, C.byref(st))
        return st

    def step(self, Ez_ext: float) -> L.PcgStats:
        st = L.PcgStats()
        L.call("hzo_step", self.h, float(Ez_ext), C.byref(st))
        return st

    def run(self, Ez) -> np.ndarray:
        Ez = np.ascontiguousarray(Ez, np.float64)
        its = np.zeros(Ez.size, np.int32)
        L.call("hzo_run", self.h, L.ptr(Ez), int(Ez.size), L.ptr(its))
        return its

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(10)
        L.call("hzo_diagnostics", self.h, L.ptr(out))
        return out

    def sweep(self, Eap_SI, delEap_SI, num_steps, path=None, max_rows=4096, max_levels=0):
        """Reference staircase P-E sweep (evolve.cu:213-410); returns the rows (n x 12)."""
        p = sweep_params(self.inp.scales, Eap_SI, delEap_SI, num_steps, max_levels)
        rows = np.zeros((max_rows, 12))
        n = C.c_int(0)
        L.call("hzo_sweep_run", self.h, C.byref(p), None if path is None else str(path).encode(), max_rows,
               L.ptr(rows), C.byref(n))
        return rows[: min(n.value, max_rows)]

    # ---------------------------------------------------------- dumps and restart
    def save_snapshot(self, pm_path, ps_path, P0_scale: float):
        """Pm and Ps as interleaved xyz f64 dumps scaled by P0 (generate.cu:266-303)."""
        L.call("hzo_save_snapshot", self.h, str(pm_path).encode(), str(ps_path).encode(), float(P0_scale))

    def checkpoint_save(self, path, peloop=0, count=0, Ez_ext=0.0, incEz=0.0):
        """Reference checkpoint file of the current P and E (generate.cu:90-168)."""
        L.call("hzo_checkpoint_save", self.h, str(path).encode(), int(peloop), int(count), float(Ez_ext),
               float(incEz))

    def checkpoint_load(self, path):
        """Restart from a reference checkpoint: P from the file, phi re-solved.
        Returns the header (peloop, count, Ez_ext, incEz) and the stored E."""
        pe, cnt = C.c_int(0), C.c_int(0)
        ez, inc = C.c_double(0.0), C.c_double(0.0)
        E = np.zeros(3 * self.inp.N)
        L.call("hzo_checkpoint_load", self.h, str(path).encode(), C.byref(pe), C.byref(cnt), C.byref(ez),
               C.byref(inc), L.ptr(E))
        return {"peloop": pe.value, "count": cnt.value, "Ez_ext": ez.value, "incEz": inc.value}, E

    def synchronize(self):
        L.call("hzo_synchronize", self.h)

    def set_timing(self, on: bool):
        L.call("hzo_set_timing", self.h, 1 if on else 0)

    def cg_counters(self):
        """Resident solves since set_timing: (solves, CG iterations, grid barriers)."""
        out = np.zeros(3, np.int64)
        L.call("hzo_cg_counters", self.h, L.ptr(out))
        return tuple(int(v) for v in out)

    def kernel_times(self):
        cnt = np.zeros(L.KCLASSES, np.int32)
        ms = np.zeros(L.KCLASSES)
        L.call("hzo_kernel_times", self.h, L.ptr(cnt), L.ptr(ms))
        return {n: (int(cnt[i]), float(ms[i])) for i, n in enumerate(L.KCLASS_NAMES)}


class Batch:
    """Films of different sizes and coefficients stepped together on one GPU
    (FP32-refined, single-reduction CG): small films share one cooperative
    resident launch per group, the others run in cache-sized groups whose
    steps replay captured hipGraphs, shared out over `lanes` host threads
    with a stream each.  The films stay owned by the caller and must outlive
    the batch."""

    def __init__(self, films, tile_kz: int = 2, use_graphs: bool = True, chunk: int = 16, lanes: int = 2):
        self.films = list(films)
        arr = (C.c_void_p * len(self.films))(*[f.h.value for f in self.films])
        self._arr = arr
        h = C.c_void_p()
        L.call("hzo_batch_create", arr, len(self.films), int(tile_kz), C.byref(h))
        self.h = h
        L.call("hzo_batch_config", h, 1 if use_graphs else 0, int(chunk))
        L.call("hzo_batch_set_lanes", h, int(lanes))

    @property
    def n(self) -> int:
        return len(self.films)

    @property
    def tile_depth(self) -> int:
        return L.lib().hzo_batch_tile_depth(self.h)

    def close(self):
        if getattr(self, "h", None):
            L.lib().hzo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, Ez, want_m: bool = False):
        """Ez: (nsteps, n) applied fields.  Returns the PCG iterations
        (nsteps, n) and, with want_m, <Pmz> of each film after each step."""
        Ez = np.ascontiguousarray(np.broadcast_to(np.asarray(Ez, np.float64), (len(Ez), self.n)))
        S = Ez.shape[0]
        its = np.zeros((S, self.n), np.int32)
        m = np.zeros((S, self.n)) if want_m else None
        L.call("hzo_batch_run", self.h, L.ptr(Ez), S, L.ptr(its), L.ptr(m))
        return (its, m) if want_m else its

    def set_lanes(self, lanes: int):
        """Host threads (1 .. 4), each with its own stream, that run the
        streaming groups side by side; the resident groups run alone."""
        L.call("hzo_batch_set_lanes", self.h, int(lanes))

    def set_shape(self, tile_kz: int, group_mib: float = 0.0):
        """Tile depth (2, 4 or 8) and group budget in MiB of CG working set (0: one group)."""
        L.call("hzo_batch_set_shape", self.h, int(tile_kz), float(group_mib))

    def film_group(self, f: int):
        """(group index, films in the group, kind) of film f; kind 0: batch
        kernels over several films, 1: the film's own kernels, 2: the
        resident solve, one cooperative launch for the group."""
        out = np.zeros(3, np.int32)
        L.call("hzo_batch_film_group", self.h, int(f), L.ptr(out))
        return int(out[0]), int(out[1]), int(out[2])

    @property
    def shape(self):
        out = np.zeros(3, np.int32)
        L.call("hzo_batch_shape", self.h, L.ptr(out))
        return {"tile_kz": int(out[0]), "group_mib": int(out[1]), "groups": int(out[2])}

    def autotune(self, Ez0, reps: int = 2):
        """Times one step per shape (tile depth 2/4/8 x group budget 0/32/64/128
        MiB, fields Ez0[f]) and keeps the fastest; returns (shape, {(kz, mib): ms})."""
        Ez0 = np.ascontiguousarray(np.broadcast_to(np.asarray(Ez0, np.float64), (self.n,)))
        kz = C.c_int(0)
        ms = np.zeros(12)
        L.call("hzo_batch_autotune", self.h, L.ptr(Ez0), int(reps), C.byref(kz), L.ptr(ms))
        times = {((2, 4, 8)[c // 4], (0, 32, 64, 128)[c % 4]): float(ms[c]) for c in range(12)}
        return self.shape, times

    def counters(self):
        out = np.zeros(2, np.int64)
        L.call("hzo_batch_counters", self.h, L.ptr(out))
        return {"cg_launches": int(out[0]), "unconverged": int(out[1])}
