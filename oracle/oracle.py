"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product library never loads it.

How each oracle function is pinned (besides the GPU comparison):
  hilbert encode/decode   reference codec compiled from its own sources
                          (oracle/_ref) -> tests/golden/hilbert_golden.npz;
                          reference known answers (test_hilbert_codec.cpp);
                          bijection / round trip / parent-child invariants
  rotdir encode           reference codec via oracle/_ref -> golden file
  sort                    numpy stable argsort (independent implementation)
  scan                    numpy cumsum
  hash build/query        Python dict + the reference's constants/capacity rule
  face neighbours         brute-force point location over all cells (tiny n)
  split / merge           reference unit-test fixtures (SplitEngineTest,
                          MergeEngineTest, AdaptationIntegrationTest) +
                          brute-force sibling search
  balance                 reference fixture (BalanceEnforcerTest); both rules
                          against an exhaustive restatement of the iteration
                          (every cell tested against every cell, flags and
                          counters compared); brute-force 2:1 checker on the
                          converged output
  key63 / cell key        prefix alignment; the packing closed form
  compact_append          closed form (numpy selection + concatenation) and
                          the split child placement rule
  merge_fields            closed form (sum in sibling order, x 0.125 rule 0)
  gen_cells / gen_flags   define the synthetic micro-benchmark input (no
                          reference counterpart); checked against the
                          workload's definition (ranges, Bernoulli rates,
                          merges only on complete sibling groups)
  lbm                     closed forms: equilibrium moments, conservation,
                          rest state, Couette profile (also across refined
                          bands), Taylor-Green decay, uniform flow through
                          coarse/fine interfaces, the momentum-exchange force
                          on a Couette wall, the pressure-outlet channel's
                          uniform steady state, y-shift invariance across a
                          periodic seam at an outlet.  The reference has no
                          lattice step, so its parity with the reference is
                          unpinned by construction.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
LIB_F32 = HERE / "_build" / "liboracle_f32.so"
REF = HERE / "_ref" / "libref_hilbert.so"

vp = C.c_void_p
sz = C.c_size_t


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if ref and Path("/root/reference").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


_lib = None
_lib_f32 = None
_ref = None


def lib():
    """The FP64 oracle (the checker)."""
    global _lib
    if _lib is None:
        _lib = _load(LIB)
    return _lib


def lib_f32():
    """The same lattice step with FP32 populations and arithmetic: timed as
    the benchmark's FP32 CPU baseline, never used as a checker."""
    global _lib_f32
    if _lib_f32 is None:
        _lib_f32 = _load(LIB_F32)
    return _lib_f32


def _load(path):
    if True:
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.or_hilbert_encode.restype = C.c_uint64
        L.or_hilbert_encode.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int]
        L.or_key63.restype = C.c_uint64
        L.or_key63.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int]
        L.or_cell_hkey.restype = C.c_uint64
        L.or_cell_hkey.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int]
        L.or_rotdir_encode.restype = C.c_uint64
        L.or_rotdir_encode.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int]
        L.or_hilbert_encode_n.argtypes = [vp, vp, vp, vp, C.c_int, vp, sz]
        L.or_hilbert_decode_n.argtypes = [vp, vp, C.c_int, vp, vp, vp, sz]
        L.or_key63_n.argtypes = [vp, vp, vp, vp, vp, sz]
        L.or_rotdir_encode_n.argtypes = [vp, vp, vp, C.c_int, vp, sz]
        L.or_sort_pairs_u64u32.argtypes = [vp, vp, sz, C.c_int, C.c_int]
        L.or_scan_exclusive_u32.restype = C.c_uint64
        L.or_scan_exclusive_u32.argtypes = [vp, vp, sz]
        L.or_hash_capacity.restype = C.c_uint64
        L.or_hash_capacity.argtypes = [sz]
        L.or_hash_build.argtypes = [vp, vp, C.c_uint64, vp, vp, sz, vp]
        L.or_hash_query.argtypes = [vp, vp, C.c_uint64, vp, vp, sz]
        L.or_face_neighbors.argtypes = [C.c_void_p, vp, vp]
        L.or_split_count.restype = sz
        L.or_split_count.argtypes = [C.c_void_p, C.c_int, vp]
        L.or_split_generate.argtypes = [C.c_void_p, vp, C.c_void_p]
        L.or_merge_mark.restype = sz
        L.or_merge_mark.argtypes = [C.c_void_p, C.c_int, vp, vp, vp]
        L.or_merge_parents.argtypes = [C.c_void_p, vp, sz, C.c_void_p]
        L.or_merge_fields.argtypes = [vp, sz, C.c_void_p, vp, C.c_int, C.c_int, vp]
        L.or_balance.restype = C.c_int
        L.or_balance.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                 C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_compact_append.restype = sz
        L.or_compact_append.argtypes = [C.c_void_p, vp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_gen_cells.argtypes = [C.c_uint64, sz, C.c_int, vp, vp, vp, vp]
        L.or_gen_flags.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]
        L.or_lbm_create.restype = vp
        L.or_lbm_create.argtypes = [C.c_void_p, vp, vp, vp, vp, vp, sz, C.c_double, vp,
                                    C.POINTER(C.c_int)]
        L.or_lbm_destroy.argtypes = [vp]
        L.or_lbm_step.argtypes = [vp, C.c_int]
        L.or_lbm_get_f.argtypes = [vp, vp, vp, vp, vp, sz, vp]
        L.or_lbm_set_f.argtypes = [vp, vp, vp, vp, vp, sz, vp]
        L.or_lbm_moments.argtypes = [vp, vp, vp, vp, vp, sz, vp, vp]
        L.or_lbm_totals.argtypes = [vp, vp, vp]
        L.or_lbm_force.argtypes = [vp, vp]
        L.or_lbm_lattice.argtypes = [vp, vp, vp]
        L.or_lbm_plant_coupling_error.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.or_lbm_vorticity.argtypes = [vp, vp, vp, vp, vp, sz, vp]
        L.or_lbm_regrid.restype = vp
        L.or_lbm_regrid.argtypes = [vp, vp, vp, vp, vp, vp, sz, C.POINTER(C.c_int)]
    return L


def ref():
    """The reference's own Hilbert codec, built from its sources (None if absent)."""
    global _ref
    if _ref is None and REF.exists():
        R = C.CDLL(str(REF))
        R.ref_hilbert_encode.restype = C.c_uint64
        R.ref_hilbert_encode.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int]
        R.ref_hilbert_decode.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        R.ref_rotdir_encode.restype = C.c_uint64
        R.ref_rotdir_encode.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int]
        R.ref_is_valid.restype = C.c_int
        R.ref_is_valid.argtypes = [C.c_uint64, C.c_int]
        _ref = R
    return _ref


def _p(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- codecs
def hilbert_encode(x, y, z, level):
    x, y, z = (np.ascontiguousarray(a, np.int32) for a in (x, y, z))
    out = np.empty(len(x), np.uint64)
    if np.isscalar(level):
        lib().or_hilbert_encode_n(_p(x), _p(y), _p(z), None, int(level), _p(out), len(x))
    else:
        lv = np.ascontiguousarray(level, np.uint8)
        lib().or_hilbert_encode_n(_p(x), _p(y), _p(z), _p(lv), 0, _p(out), len(x))
    return out


def hilbert_decode(h, level):
    h = np.ascontiguousarray(h, np.uint64)
    x, y, z = (np.empty(len(h), np.int32) for _ in range(3))
    if np.isscalar(level):
        lib().or_hilbert_decode_n(_p(h), None, int(level), _p(x), _p(y), _p(z), len(h))
    else:
        lv = np.ascontiguousarray(level, np.uint8)
        lib().or_hilbert_decode_n(_p(h), _p(lv), 0, _p(x), _p(y), _p(z), len(h))
    return x, y, z


def key63(x, y, z, level):
    x, y, z = (np.ascontiguousarray(a, np.int32) for a in (x, y, z))
    lv = np.ascontiguousarray(level, np.uint8)
    out = np.empty(len(x), np.uint64)
    lib().or_key63_n(_p(x), _p(y), _p(z), _p(lv), _p(out), len(x))
    return out


def cell_hkey(x, y, z, level):
    return np.array([lib().or_cell_hkey(int(a), int(b), int(c), int(l))
                     for a, b, c, l in zip(x, y, z, level)], np.uint64)


def rotdir_encode(x, y, z, max_level):
    x, y, z = (np.ascontiguousarray(a, np.int32) for a in (x, y, z))
    out = np.empty(len(x), np.uint64)
    lib().or_rotdir_encode_n(_p(x), _p(y), _p(z), int(max_level), _p(out), len(x))
    return out


# ---------------------------------------------------------------- sort/scan/hash
def sort_pairs(keys, vals, begin_bit=0, end_bit=64):
    k = np.array(keys, np.uint64)
    v = np.array(vals, np.uint32)
    lib().or_sort_pairs_u64u32(_p(k), _p(v), len(k), begin_bit, end_bit)
    return k, v


def scan_exclusive(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    tot = lib().or_scan_exclusive_u32(_p(a), _p(out), len(a))
    return out, int(tot)


def hash_capacity(n):
    return int(lib().or_hash_capacity(n))


def hash_build(keys, vals, cap):
    keys = np.ascontiguousarray(keys, np.uint64)
    vals = np.ascontiguousarray(vals, np.uint32)
    tk = np.empty(cap, np.uint64)
    tv = np.empty(cap, np.uint32)
    pc = np.empty(len(keys), np.uint32)
    lib().or_hash_build(_p(tk), _p(tv), cap, _p(keys), _p(vals), len(keys), _p(pc))
    return tk, tv, pc


def hash_query(tk, tv, cap, q):
    q = np.ascontiguousarray(q, np.uint64)
    out = np.empty(len(q), np.uint32)
    lib().or_hash_query(_p(tk), _p(tv), cap, _p(q), _p(out), len(q))
    return out


# ---------------------------------------------------------------- AMR
class _Cells(C.Structure):
    _fields_ = [("x", vp), ("y", vp), ("z", vp), ("level", vp), ("state", vp), ("flag", vp),
                ("material", vp), ("n", sz)]


class CellSet:
    """Host SoA cell arrays (native coords) mirrored into the oracle struct."""

    FIELDS = (("x", np.int32), ("y", np.int32), ("z", np.int32), ("level", np.uint8),
              ("state", np.uint8), ("flag", np.int32), ("material", np.uint32))

    def __init__(self, n=0, **arrays):
        if not n and arrays.get("x") is not None:
            n = len(arrays["x"])  # missing fields get zeros of the given arrays' length
        for name, dt in self.FIELDS:
            a = arrays.get(name)
            setattr(self, name, np.zeros(n, dt) if a is None else np.ascontiguousarray(a, dt))
        self.n = len(self.x)

    def struct(self):
        return _Cells(*(getattr(self, nm).ctypes.data for nm, _ in self.FIELDS), self.n)

    def copy(self):
        return CellSet(**{nm: getattr(self, nm).copy() for nm, _ in self.FIELDS})

    def head(self, n):
        return CellSet(**{nm: getattr(self, nm)[:n].copy() for nm, _ in self.FIELDS})


def _root(root):
    return np.ascontiguousarray(root if root is not None else (1, 1, 1), np.int32)


def face_neighbors(c: CellSet, root=None):
    r = _root(root)
    out = np.empty(6 * c.n, np.uint32)
    s = c.struct()
    lib().or_face_neighbors(C.byref(s), _p(r), _p(out))
    return out


def split(c: CellSet, max_level):
    cs = np.empty(c.n, np.uint32)
    s = c.struct()
    ns = lib().or_split_count(C.byref(s), max_level, _p(cs))
    ch = CellSet(8 * ns)
    chs = ch.struct()
    lib().or_split_generate(C.byref(s), _p(cs), C.byref(chs))
    return cs, ch


def merge(c: CellSet, check_balance=0, root=None):
    r = _root(root)
    gid = np.empty(c.n, np.uint32)
    lead = np.empty(max(c.n, 1), np.uint32)
    s = c.struct()
    ng = lib().or_merge_mark(C.byref(s), check_balance, _p(r), _p(gid), _p(lead))
    par = CellSet(ng)
    ps = par.struct()
    lib().or_merge_parents(C.byref(s), _p(lead), ng, C.byref(ps))
    return gid, lead[:ng].copy(), par


def merge_fields(c: CellSet, leaders, fields, ncomp, rule):
    fields = np.ascontiguousarray(fields, np.float32)
    out = np.empty(len(leaders) * ncomp, np.float32)
    s = c.struct()
    lead = np.ascontiguousarray(leaders, np.uint32)
    lib().or_merge_fields(_p(lead), len(lead), C.byref(s), _p(fields), ncomp, rule, _p(out))
    return out


def balance(c: CellSet, mode, anchor_level, max_split_level, max_iters=64, root=None):
    """Runs in place on c.flag; returns (iterations, converged, violations, marked)."""
    r = _root(root)
    conv = C.c_int(0)
    tv = C.c_uint32(0)
    tm = C.c_uint32(0)
    s = c.struct()
    it = lib().or_balance(C.byref(s), mode, anchor_level, max_split_level, max_iters, _p(r),
                          C.byref(conv), C.byref(tv), C.byref(tm))
    return it, conv.value, tv.value, tm.value


def compact_append(c: CellSet, gid, children: CellSet, parents: CellSet):
    out = CellSet(c.n + children.n + parents.n)
    s, cs, ps, os_ = c.struct(), children.struct(), parents.struct(), out.struct()
    n = lib().or_compact_append(C.byref(s), _p(np.ascontiguousarray(gid, np.uint32)),
                                C.byref(cs), C.byref(ps), C.byref(os_))
    return out.head(n)


def regrid(c: CellSet, balance_mode, anchor_level, max_split_level, max_iters=64, root=None,
           merge_balance_check=0):
    """Balance -> merge -> compaction -> split -> parents, as fl_regrid orders them."""
    c = c.copy()
    res = None
    if balance_mode >= 0:
        res = balance(c, balance_mode, anchor_level, max_split_level, max_iters, root)
    gid, lead, par = merge(c, merge_balance_check, root)
    _, ch = split(c, max_split_level)
    return compact_append(c, gid, ch, par), res, gid, lead


def _thresh(p):
    return (1 << 64) - 1 if p >= 1.0 else int(np.ldexp(p, 64))


def gen_flags(c: CellSet, seed, p_split, p_merge):
    """In place: the micro-benchmark's split / merge flags (as fl_gen_flags)."""
    s = c.struct()
    lib().or_gen_flags(C.byref(s), seed, _thresh(p_split), _thresh(p_merge))
    return c


def gen_cells(seed, n, max_level):
    x, y, z = (np.empty(n, np.int32) for _ in range(3))
    lv = np.empty(n, np.uint8)
    lib().or_gen_cells(seed, n, max_level, _p(x), _p(y), _p(z), _p(lv))
    return x, y, z, lv


# ---------------------------------------------------------------- LBM
class _LbmParams(C.Structure):
    _fields_ = [("nx0", C.c_int), ("ny0", C.c_int), ("nz0", C.c_int), ("max_level", C.c_int),
                ("tau0", C.c_double), ("bc", C.c_int * 6), ("bc_u", (C.c_double * 3) * 6),
                ("rho0", C.c_double)]


def lattice():
    c = np.empty(57, np.int32)
    w = np.empty(19, np.float64)
    o = np.empty(19, np.int32)
    lib().or_lbm_lattice(_p(c), _p(w), _p(o))
    return c.reshape(19, 3), w, o


class Lbm:
    """FP64 oracle solver over a leaf list (dense per-level arrays)."""

    def __init__(self, nx0, ny0, nz0, max_level, tau0, bc, bc_u, x, y, z, level, state,
                 rho_init=1.0, u_init=(0.0, 0.0, 0.0), rho0=1.0, single=False):
        self.L = lib_f32() if single else lib()
        p = _LbmParams()
        p.nx0, p.ny0, p.nz0, p.max_level, p.tau0, p.rho0 = nx0, ny0, nz0, max_level, tau0, rho0
        for f in range(6):
            p.bc[f] = int(bc[f])
            for a in range(3):
                p.bc_u[f][a] = float(bc_u[f][a])
        self.x, self.y, self.z = (np.ascontiguousarray(a, np.int32) for a in (x, y, z))
        self.level = np.ascontiguousarray(level, np.uint8)
        self.state = np.ascontiguousarray(state, np.uint8)
        u = np.array(u_init, np.float64)
        err = C.c_int(0)
        self.h = self.L.or_lbm_create(C.byref(p), _p(self.x), _p(self.y), _p(self.z), _p(self.level),
                                     _p(self.state), len(self.x), float(rho_init), _p(u),
                                     C.byref(err))
        if not self.h:
            raise ValueError(f"oracle mesh error {err.value}")
        fl = self.state == 0
        self.fx, self.fy, self.fz, self.flv = self.x[fl], self.y[fl], self.z[fl], self.level[fl]

    def step(self, n=1):
        self.L.or_lbm_step(self.h, n)

    def plant_coupling_error(self, c2f_mult, f2c_mult, slopes=1.0):
        self.L.or_lbm_plant_coupling_error(self.h, c2f_mult, f2c_mult, slopes)

    def get_f(self):
        out = np.empty((len(self.fx), 19), np.float64)
        self.L.or_lbm_get_f(self.h, _p(self.fx), _p(self.fy), _p(self.fz), _p(self.flv),
                           len(self.fx), _p(out))
        return out

    def set_f(self, f):
        f = np.ascontiguousarray(f, np.float64)
        self.L.or_lbm_set_f(self.h, _p(self.fx), _p(self.fy), _p(self.fz), _p(self.flv),
                           len(self.fx), _p(f))

    def moments(self):
        rho = np.empty(len(self.fx))
        u = np.empty((len(self.fx), 3))
        self.L.or_lbm_moments(self.h, _p(self.fx), _p(self.fy), _p(self.fz), _p(self.flv),
                             len(self.fx), _p(rho), _p(u))
        return rho, u

    def totals(self):
        m = np.zeros(1)
        j = np.zeros(3)
        self.L.or_lbm_totals(self.h, _p(m), _p(j))
        return float(m[0]), j

    def force(self):
        F = np.zeros(3)
        self.L.or_lbm_force(self.h, _p(F))
        return F

    def vorticity(self):
        out = np.empty(len(self.fx))
        self.L.or_lbm_vorticity(self.h, _p(self.fx), _p(self.fy), _p(self.fz), _p(self.flv), len(self.fx),
                               _p(out))
        return out

    def regrid(self, x, y, z, level, state):
        """Replace the mesh by a new leaf list, carrying the populations over."""
        x, y, z = (np.ascontiguousarray(a, np.int32) for a in (x, y, z))
        level = np.ascontiguousarray(level, np.uint8)
        state = np.ascontiguousarray(state, np.uint8)
        err = C.c_int(0)
        h = self.L.or_lbm_regrid(self.h, _p(x), _p(y), _p(z), _p(level), _p(state), len(x), C.byref(err))
        if not h:
            raise ValueError(f"oracle mesh error {err.value}")
        self.L.or_lbm_destroy(self.h)
        self.h = h
        self.x, self.y, self.z, self.level, self.state = x, y, z, level, state
        fl = state == 0
        self.fx, self.fy, self.fz, self.flv = x[fl], y[fl], z[fl], level[fl]

    def __del__(self):
        try:
            if self.h:
                self.L.or_lbm_destroy(self.h)
                self.h = None
        except Exception:
            pass
