"""Python binding of libphc.so (ctypes over the C ABI in include/phc.h).

The classes mirror the reference's solver surface so that tests and the
benchmark read like the reference's own drivers:
  FdSolver      FDTDKernel::addRHS / computeOneStepSB (src/fd/utilsFDTD/
                FDTDkernels.hpp:4-131) + FDTDUtils::output (FDTDutils.hpp:193)
  SemSolver     SEMsolver::computeOneStep / outputPnValues
                (src/sem/solver/SEMsolver.cpp:21-142)
  fd_setup      FDTDInit::init_coefficients/init_source/init_models
                (src/fd/utilsFDTD/FDTDinit.hpp:202-317)

There is no fallback: if libphc.so is missing or a call fails, this module
raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphc.so")
_lib = None

ARITH_REF = 0
ARITH_FP32 = 1
EXP_DOUBLE = 0
EXP_FLOAT = 1

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
c_int, c_float, c_double, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p


class PhcError(RuntimeError):
    pass


class FdParams(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int),
                ("lx", c_int), ("ly", c_int), ("lz", c_int),
                ("xs", c_int), ("ys", c_int), ("zs", c_int),
                ("coefx", c_void_p), ("coefy", c_void_p), ("coefz", c_void_p),
                ("coef0", c_double),
                ("vp", c_void_p), ("rhs", c_void_p), ("nsamples", c_int),
                ("sponge_width", c_int), ("sponge_alpha", c_float),
                ("init_value", c_float), ("arith", c_int),
                ("tile_j", c_int), ("block_waves", c_int), ("chunk_i", c_int),
                ("index64", c_int), ("k_remainder", c_int)]


# name -> (restype, argtypes) of every exported C symbol declared in include/*.h
SIGNATURES = {
    "phc_last_error": (ctypes.c_char_p, []),
    "phc_version": (ctypes.c_char_p, []),
    "phc_device_count": (c_int, []),
    "phc_device_synchronize": (c_int, []),
    "phc_copy_bandwidth": (c_int, [ctypes.c_size_t, c_int, ctypes.POINTER(c_double)]),
    # setup (include/phc_setup.h)
    "phc_fd_init_coef": (c_int, [c_int, c_float, f32p]),
    "phc_fd_coef0": (c_double, [f32p, c_int, f32p, c_int, f32p, c_int]),
    "phc_fd_compute_dt": (c_float, [c_float, f32p, c_int, f32p, c_int, f32p, c_int, c_float]),
    "phc_ricker": (c_float, [c_float, c_float, c_int, c_int]),
    "phc_source_term": (c_int, [c_int, c_float, c_float, c_int, c_int, f32p]),
    "phc_fd_model_two_layer": (c_int, [c_int, c_int, c_int, c_float, c_float, f32p]),
    "phc_fd_model_layered": (c_int, [c_int, c_int, c_int, c_int, c_float, c_float,
                                     ctypes.c_uint64, c_float, c_void_p, f32p]),
    "phc_sem_model_gaussian": (c_int, [c_int, c_int, c_int, c_double, c_float, c_float,
                                       ctypes.c_uint64, f32p]),
    # FD
    "phc_fd_create": (c_int, [ctypes.POINTER(FdParams), ctypes.POINTER(c_void_p)]),
    "phc_fd_destroy": (c_int, [c_void_p]),
    "phc_fd_add_rhs": (c_int, [c_void_p, c_int, c_int]),
    "phc_fd_add_rhs_value": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_float]),
    "phc_fd_compute_one_step_sb": (c_int, [c_void_p, c_int, c_int]),
    "phc_fd_run": (c_int, [c_void_p, c_int, c_int, ctypes.POINTER(c_int)]),
    "phc_fd_set_receivers": (c_int, [c_void_p, c_int, c_void_p]),
    "phc_fd_get_traces": (c_int, [c_void_p, c_void_p, c_void_p]),
    "phc_fd_get_field": (c_int, [c_void_p, c_int, f32p]),
    "phc_fd_set_field": (c_int, [c_void_p, c_int, f32p]),
    "phc_fd_read_point": (c_int, [c_void_p, c_int, c_int, c_int, c_int, ctypes.POINTER(c_float)]),
    "phc_fd_get_slice_xz": (c_int, [c_void_p, c_int, c_int, f32p]),
    "phc_fd_check_finite": (c_int, [c_void_p, c_int]),
    "phc_fd_last_run_kernel_ms": (c_int, [c_void_p, ctypes.POINTER(c_double)]),
    "phc_fd_set_timing": (c_int, [c_void_p, c_int]),
    "phc_fd_synchronize": (c_int, [c_void_p]),
    "phc_fd_autotune": (c_int, [c_void_p, c_int, ctypes.c_char_p, ctypes.POINTER(c_int),
                                ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "phc_fd_set_pml": (c_int, [c_void_p, i32p, f32p, f32p, f32p, f32p]),
    "phc_fd_compute_one_step_pml": (c_int, [c_void_p, c_int, c_int]),
    "phc_fd_pml_width": (c_int, [c_float, c_float, c_float]),
    "phc_fd_pml_scale": (c_float, [c_float, c_float, c_int, c_float]),
    "phc_fd_pml_profile": (c_int, [c_int, c_int, c_float, f32p]),
}


def lib():
    """Load libphc.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PhcError("libphc.so not found at %s: run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names from SIGNATURES that libphc.so actually exports."""
    L = lib()
    return [n for n in SIGNATURES if hasattr(L, n)]


def check(st):
    if st != 0:
        raise PhcError("libphc error %d: %s" % (st, lib().phc_last_error().decode()))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def device_count():
    return lib().phc_device_count()


# ------------------------------------------------------------------ setup
def fd_coefs(L, dx):
    c = np.zeros(L + 1, np.float32)
    check(lib().phc_fd_init_coef(L, dx, c))
    return c


def source_term(nsamples, dt, f0, order=1, exp_variant=EXP_DOUBLE):
    out = np.zeros(nsamples, np.float32)
    check(lib().phc_source_term(nsamples, dt, f0, order, exp_variant, out))
    return out


class FdSetup:
    """FDTDInit (src/fd/utilsFDTD/FDTDinit.hpp:36-317) for one grid: coefficients,
    dt, source wavelet and model.  model: "two_layer" (the reference's
    init_models), "homogeneous" (vmin everywhere), ("layered", nlayers, seed)
    or an explicit float32 array of vp^2 dt^2 (IDX3 order)."""

    def __init__(self, nx=150, ny=None, nz=None, L=4, dx=10.0, xs=None, ys=None, zs=None,
                 f0=15.0, time_max=1.0, vmin=1500.0, vmax=4500.0, cfl=0.8, source_order=1,
                 sponge_width=20, sponge_alpha=-0.00015, model="two_layer", nsamples=None,
                 ly=None, lz=None):
        self.nx = nx
        self.ny = nx if ny is None else ny
        self.nz = nx if nz is None else nz
        # half widths per axis (-lx / -ly / -lz, FDTDinit.hpp:96-110); L is lx
        self.L = L
        self.lx = L
        self.ly = L if ly is None else ly
        self.lz = L if lz is None else lz
        self.xs = nx // 2 if xs is None else xs
        self.ys = self.ny // 2 if ys is None else ys
        self.zs = self.nz // 2 if zs is None else zs
        self.cx = fd_coefs(self.lx, dx)
        self.cy = fd_coefs(self.ly, dx)
        self.cz = fd_coefs(self.lz, dx)
        self.coef0 = lib().phc_fd_coef0(self.cx, self.lx, self.cy, self.ly, self.cz, self.lz)
        self.dt = np.float32(lib().phc_fd_compute_dt(vmax, self.cx, self.lx, self.cy, self.ly, self.cz,
                                                     self.lz, cfl))
        self.nsamples = int(np.float32(time_max) / self.dt) if nsamples is None else nsamples
        self.f0 = f0
        self.rhs = source_term(self.nsamples, self.dt, f0, source_order, EXP_DOUBLE)
        n3 = self.nx * self.ny * self.nz
        self.vp = np.zeros(n3, np.float32)
        if isinstance(model, np.ndarray):
            if model.size != n3:
                raise ValueError("explicit vp^2 dt^2 model has %d values, grid has %d" % (model.size, n3))
            self.vp = np.ascontiguousarray(model, np.float32).reshape(-1).copy()
        elif model == "two_layer":
            check(lib().phc_fd_model_two_layer(self.nx, self.ny, self.nz, vmin, self.dt, self.vp))
        elif model == "homogeneous":
            v0 = np.float32(np.float32(np.float32(vmin) * np.float32(vmin)) * self.dt) * self.dt
            self.vp[:] = v0
        elif isinstance(model, tuple) and model[0] == "layered":
            nl, seed = model[1], model[2]
            self.interfaces = np.zeros(max(nl - 1, 1), np.int32)
            check(lib().phc_fd_model_layered(self.nx, self.ny, self.nz, nl, vmin, vmax, seed,
                                             self.dt, _ptr(self.interfaces), self.vp))
        else:
            raise ValueError(model)
        self.sponge_width = sponge_width
        self.sponge_alpha = sponge_alpha


class FdSolver:
    """Device-resident FD state: two levels (pnGlobal(:,0), pnGlobal(:,1)) plus
    vp^2 dt^2, sponge and source, with the FDTDKernel step surface."""

    def __init__(self, setup, arith=ARITH_REF, init_value=1e-6, tile_j=0, chunk_i=0,
                 index64=False, block_waves=0, k_remainder=0):
        s = setup
        self.setup = s
        self.nx, self.ny, self.nz, self.L = s.nx, s.ny, s.nz, s.L
        self.lh = (s.lx, s.ly, s.lz)
        self._keep = [s.cx, s.cy, s.cz, s.vp, s.rhs]
        p = FdParams()
        p.nx, p.ny, p.nz = s.nx, s.ny, s.nz
        p.lx, p.ly, p.lz = s.lx, s.ly, s.lz
        p.xs, p.ys, p.zs = s.xs, s.ys, s.zs
        p.coefx, p.coefy, p.coefz = _ptr(s.cx), _ptr(s.cy), _ptr(s.cz)
        p.coef0 = s.coef0
        p.vp = _ptr(s.vp)
        p.rhs = _ptr(s.rhs)
        p.nsamples = s.nsamples
        p.sponge_width = s.sponge_width
        p.sponge_alpha = s.sponge_alpha
        p.init_value = init_value
        p.arith = arith
        p.tile_j = tile_j
        p.block_waves = block_waves
        p.chunk_i = chunk_i
        p.index64 = 1 if index64 else 0
        p.k_remainder = k_remainder
        h = c_void_p()
        check(lib().phc_fd_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.init_value = init_value
        self.i1, self.i2 = 0, 1   # FDTDInit::i1/i2 (FDTDinit.hpp:28-29)
        self.nrec = 0

    def close(self):
        if self.h:
            check(lib().phc_fd_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # FDTDKernel surface
    def add_rhs(self, it, cb):
        check(lib().phc_fd_add_rhs(self.h, it, cb))

    def compute_one_step_sb(self, ca, cb):
        check(lib().phc_fd_compute_one_step_sb(self.h, ca, cb))

    def compute_one_step_pml(self, ca, cb):
        check(lib().phc_fd_compute_one_step_pml(self.h, ca, cb))

    def set_pml(self, ndamp=None, f0=None, vmin=1500.0, vmax=4500.0, dx=10.0):
        """FDTDInit -usePML: width (int)(4 vmin/(2.5 f0)/dx) unless given,
        init_eta profiles from the setup's dt; ndamp=0 removes the PML."""
        s = self.setup
        f0 = s.f0 if f0 is None else f0
        nd = lib().phc_fd_pml_width(vmin, f0, dx) if ndamp is None else ndamp
        prof = []
        for n in (self.nx, self.ny, self.nz):
            p = np.zeros(n + 2, np.float32)
            if nd > 0:
                check(lib().phc_fd_pml_profile(n, nd, lib().phc_fd_pml_scale(s.dt, vmax, nd, dx), p))
            prof.append(p)
        hd2 = np.full(3, np.float32(1.0 / (4.0 * dx * dx)), np.float32)
        check(lib().phc_fd_set_pml(self.h, np.array([nd] * 3, np.int32), prof[0], prof[1], prof[2], hd2))
        self.pml = nd > 0
        self.ndamp = nd
        return nd

    def step(self, it):
        """One iteration of FDTDmain.cpp:55-81 through the per-call surface."""
        self.add_rhs(it, self.i2)
        if getattr(self, "pml", False):
            self.compute_one_step_pml(self.i1, self.i2)
        else:
            self.compute_one_step_sb(self.i1, self.i2)
        self.i1, self.i2 = self.i2, self.i1

    def run(self, it0, nsteps):
        """Fused, graph-captured loop; returns (source trace, receiver traces)."""
        cur = c_int(self.i2)
        check(lib().phc_fd_run(self.h, it0, nsteps, ctypes.byref(cur)))
        self.i2 = cur.value
        self.i1 = 1 - cur.value
        src = np.zeros(nsteps, np.float32)
        rec = np.zeros(max(self.nrec, 1) * nsteps, np.float32)
        check(lib().phc_fd_get_traces(self.h, _ptr(src), _ptr(rec) if self.nrec else None))
        return src, (rec.reshape(self.nrec, nsteps) if self.nrec else None)

    def set_receivers(self, ijk):
        a = np.ascontiguousarray(ijk, np.int32).reshape(-1)
        self.nrec = len(a) // 3
        check(lib().phc_fd_set_receivers(self.h, self.nrec, _ptr(a)))

    def field(self, level):
        lx, ly, lz = self.lh
        shape = (self.nx + 2 * lx, self.ny + 2 * ly, self.nz + 2 * lz)
        out = np.zeros(shape[0] * shape[1] * shape[2], np.float32)
        check(lib().phc_fd_get_field(self.h, level, out))
        return out.reshape(shape)

    def set_field(self, level, arr):
        check(lib().phc_fd_set_field(self.h, level, np.ascontiguousarray(arr, np.float32).reshape(-1)))

    def current(self):
        return self.field(self.i2)

    def previous(self):
        return self.field(self.i1)

    def read_point(self, level, i, j, k):
        v = c_float()
        check(lib().phc_fd_read_point(self.h, level, i, j, k, ctypes.byref(v)))
        return v.value

    def slice_xz(self, level, j):
        out = np.zeros(self.nz * self.nx, np.float32)
        check(lib().phc_fd_get_slice_xz(self.h, level, j, out))
        return out.reshape(self.nz, self.nx)

    def check_finite(self, level):
        return lib().phc_fd_check_finite(self.h, level)

    def set_timing(self, on):
        check(lib().phc_fd_set_timing(self.h, 1 if on else 0))

    def last_kernel_ms(self):
        v = c_double()
        check(lib().phc_fd_last_run_kernel_ms(self.h, ctypes.byref(v)))
        return v.value

    def synchronize(self):
        check(lib().phc_fd_synchronize(self.h))

    def set_learned_boundary(self, lb, halo=None, layer_damping=0.0):
        """Attach a LearnedBoundary (None detaches); halo defaults to init_value."""
        h = self.init_value if halo is None else halo
        check(lib().phc_fd_set_learned_boundary(self.h, lb.h if lb is not None else None, h, layer_damping))

    def lb_corrections(self, cur=None):
        """Per-face corrections the next step would add (list of 6 arrays)."""
        sizes = lb_corr_sizes(self.nx, self.ny, self.nz)
        out = np.zeros(sum(sizes), np.float32)
        check(lib().phc_fd_lb_corrections(self.h, self.i2 if cur is None else cur, out))
        return np.split(out, np.cumsum(sizes)[:-1])

    def autotune(self, steps_per_trial=10, cache_path=None):
        """Pick the fastest block shape for this grid; returns (tile_j, waves, chunk_i)."""
        tj, ty, ci = c_int(), c_int(), c_int()
        cp = None if cache_path is None else cache_path.encode()
        check(lib().phc_fd_autotune(self.h, steps_per_trial, cp, ctypes.byref(tj), ctypes.byref(ty),
                                    ctypes.byref(ci)))
        return tj.value, ty.value, ci.value


# ------------------------------------------------------------------- FD64
SIGNATURES.update({
    "phc_fd64_create": (c_int, [ctypes.POINTER(FdParams), ctypes.POINTER(c_void_p)]),
    "phc_fd64_destroy": (c_int, [c_void_p]),
    "phc_fd64_set_receivers": (c_int, [c_void_p, c_int, c_void_p]),
    "phc_fd64_run": (c_int, [c_void_p, c_int, c_int]),
    "phc_fd64_get_traces": (c_int, [c_void_p, c_void_p, c_void_p]),
    "phc_fd64_get_field": (c_int, [c_void_p, c_int, f64p]),
})


class Fd64Solver:
    """The FD loop in FP64 (accuracy yardstick), from an FdSetup."""

    def __init__(self, setup, init_value=1e-6):
        s = setup
        self.setup = s
        self.nx, self.ny, self.nz, self.L = s.nx, s.ny, s.nz, s.L
        p = FdParams()
        p.nx, p.ny, p.nz = s.nx, s.ny, s.nz
        p.lx = p.ly = p.lz = s.L
        p.xs, p.ys, p.zs = s.xs, s.ys, s.zs
        p.coefx, p.coefy, p.coefz = _ptr(s.cx), _ptr(s.cy), _ptr(s.cz)
        p.coef0 = s.coef0
        p.vp, p.rhs, p.nsamples = _ptr(s.vp), _ptr(s.rhs), s.nsamples
        p.sponge_width, p.sponge_alpha, p.init_value = s.sponge_width, s.sponge_alpha, init_value
        h = c_void_p()
        check(lib().phc_fd64_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.it = 0
        self.nrec = 0

    def close(self):
        if self.h:
            check(lib().phc_fd64_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_receivers(self, ijk):
        a = np.ascontiguousarray(ijk, np.int32).reshape(-1)
        self.nrec = len(a) // 3
        check(lib().phc_fd64_set_receivers(self.h, self.nrec, _ptr(a)))

    def run(self, nsteps):
        check(lib().phc_fd64_run(self.h, self.it, nsteps))
        self.it += nsteps
        src = np.zeros(nsteps)
        rec = np.zeros(max(self.nrec, 1) * nsteps)
        check(lib().phc_fd64_get_traces(self.h, _ptr(src), _ptr(rec) if self.nrec else None))
        return src, (rec.reshape(self.nrec, nsteps) if self.nrec else None)

    def field(self, current=True):
        L = self.L
        out = np.zeros((self.nx + 2 * L) * (self.ny + 2 * L) * (self.nz + 2 * L))
        check(lib().phc_fd64_get_field(self.h, 1 if current else 0, out))
        return out.reshape(self.nx + 2 * L, self.ny + 2 * L, self.nz + 2 * L)


# ------------------------------------------------------------- batched shots
class BatchParams(ctypes.Structure):
    _fields_ = [("nshots", c_int), ("dims", c_void_p), ("src", c_void_p),
                ("vp", c_void_p), ("rhs", c_void_p), ("nsamples", c_int),
                ("half_width", c_int), ("coef", c_void_p), ("coef0", c_double),
                ("sponge_width", c_int), ("sponge_alpha", c_float), ("init_value", c_float),
                ("arith", c_int), ("tile_j", c_int), ("block_waves", c_int),
                ("chunk_i", c_int), ("blocks_per_cu", c_int), ("k_remainder", c_int)]


SIGNATURES.update({
    "phc_batch_create": (c_int, [ctypes.POINTER(BatchParams), ctypes.POINTER(c_void_p)]),
    "phc_batch_destroy": (c_int, [c_void_p]),
    "phc_batch_set_receivers": (c_int, [c_void_p, c_int, c_void_p]),
    "phc_batch_run": (c_int, [c_void_p, c_int, c_int]),
    "phc_batch_get_traces": (c_int, [c_void_p, c_void_p, c_void_p]),
    "phc_batch_get_field": (c_int, [c_void_p, c_int, c_int, f32p]),
    "phc_batch_set_timing": (c_int, [c_void_p, c_int]),
    "phc_batch_last_run_kernel_ms": (c_int, [c_void_p, ctypes.POINTER(c_double)]),
    "phc_batch_synchronize": (c_int, [c_void_p]),
    "phc_batch_autotune": (c_int, [c_void_p, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                   ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
})


def multi_shot_setups(nshots=64, nmin=128, nmax=384, f0_range=(5.0, 15.0), nsteps=800, seed=2024,
                      nlayers=6, sponge_width=20, cfl=0.4, dx=10.0):
    """The batched survey: per shot, each axis drawn uniformly from [nmin, nmax],
    a centred source, a layered vp model of its own seed and a Ricker f0 drawn
    from f0_range; shared h, dt (from vmax and the CFL number) and step count."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nshots):
        nx, ny, nz = (int(v) for v in rng.integers(nmin, nmax + 1, size=3))
        f0 = float(rng.uniform(*f0_range))
        out.append(FdSetup(nx, ny, nz, L=4, dx=dx, f0=f0, cfl=cfl, sponge_width=sponge_width,
                           model=("layered", nlayers, seed * 1000 + s), nsamples=nsteps))
    return out


class FdBatch:
    """Many FD shots (FdSetup each, same half width, coefficients, sponge and
    sample count) advanced together by the batched persistent kernel."""

    def __init__(self, setups, arith=ARITH_FP32, init_value=1e-6, tile_j=0, block_waves=0, chunk_i=0,
                 blocks_per_cu=0, k_remainder=0):
        s0 = setups[0]
        for s in setups:
            if (s.L != s0.L or s.nsamples != s0.nsamples or s.sponge_width != s0.sponge_width
                    or not np.array_equal(s.cx, s0.cx) or not np.array_equal(s.cy, s0.cx)
                    or not np.array_equal(s.cz, s0.cx)):
                raise ValueError("batched shots must share half width, coefficients, sponge and nsamples")
        self.setups = setups
        self.n = len(setups)
        self.dims = np.array([[s.nx, s.ny, s.nz] for s in setups], np.int32)
        src = np.array([[s.xs, s.ys, s.zs] for s in setups], np.int32)
        vps = (c_void_p * self.n)(*[_ptr(s.vp) for s in setups])
        rhs = (c_void_p * self.n)(*[_ptr(s.rhs) for s in setups])
        p = BatchParams()
        p.nshots = self.n
        p.dims = _ptr(self.dims)
        p.src = _ptr(src)
        p.vp = ctypes.cast(vps, c_void_p)
        p.rhs = ctypes.cast(rhs, c_void_p)
        p.nsamples = s0.nsamples
        p.half_width = s0.L
        p.coef = _ptr(s0.cx)
        p.coef0 = s0.coef0
        p.sponge_width = s0.sponge_width
        p.sponge_alpha = s0.sponge_alpha
        p.init_value = init_value
        p.arith = arith
        p.tile_j, p.block_waves, p.chunk_i, p.blocks_per_cu = tile_j, block_waves, chunk_i, blocks_per_cu
        p.k_remainder = k_remainder
        h = c_void_p()
        check(lib().phc_batch_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.it = 0
        self.nrec = 0

    def close(self):
        if self.h:
            check(lib().phc_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_receivers(self, ijk):
        """ijk: (nshots, nrec, 3) grid points."""
        a = np.ascontiguousarray(ijk, np.int32)
        self.nrec = a.shape[1]
        check(lib().phc_batch_set_receivers(self.h, self.nrec, _ptr(a.reshape(-1))))

    def run(self, nsteps, traces=True, it0=None):
        """Advance every shot by nsteps samples (from sample it0, default: where
        the last run ended); returns (source traces (nshots, nsteps), receiver
        traces (nshots, nrec, nsteps) or None)."""
        it0 = self.it if it0 is None else it0
        check(lib().phc_batch_run(self.h, it0, nsteps))
        self.it = it0 + nsteps
        if not traces:
            return None, None
        src = np.zeros(self.n * nsteps, np.float32)
        rec = np.zeros(max(self.nrec, 1) * self.n * nsteps, np.float32)
        check(lib().phc_batch_get_traces(self.h, _ptr(src), _ptr(rec) if self.nrec else None))
        return (src.reshape(self.n, nsteps),
                rec.reshape(self.n, self.nrec, nsteps) if self.nrec else None)

    def field(self, shot, current=True):
        s = self.setups[shot]
        L = s.L
        out = np.zeros((s.nx + 2 * L) * (s.ny + 2 * L) * (s.nz + 2 * L), np.float32)
        check(lib().phc_batch_get_field(self.h, shot, 1 if current else 0, out))
        return out.reshape(s.nx + 2 * L, s.ny + 2 * L, s.nz + 2 * L)

    def set_timing(self, on):
        check(lib().phc_batch_set_timing(self.h, 1 if on else 0))

    def last_kernel_ms(self):
        v = c_double()
        check(lib().phc_batch_last_run_kernel_ms(self.h, ctypes.byref(v)))
        return v.value

    def synchronize(self):
        check(lib().phc_batch_synchronize(self.h))

    def autotune(self, steps_per_trial=3):
        """Returns (tile_j, block_waves, chunk_i, blocks_per_cu) of the fastest trial."""
        r = [c_int() for _ in range(4)]
        check(lib().phc_batch_autotune(self.h, steps_per_trial, *[ctypes.byref(x) for x in r]))
        return tuple(x.value for x in r)

    def points(self):
        return int(np.prod(self.dims, axis=1, dtype=np.int64).sum())


# --------------------------------------------------------- learned boundary
LB_WIDTH = 4
LB_FACES = 6
LB_NP = 1825
LB_SAMPLE = 22
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

SIGNATURES.update({
    "phc_lb_create": (c_int, [ctypes.c_uint64, ctypes.POINTER(c_void_p)]),
    "phc_lb_destroy": (c_int, [c_void_p]),
    "phc_lb_get_params": (c_int, [c_void_p, f32p]),
    "phc_lb_set_params": (c_int, [c_void_p, f32p]),
    "phc_lb_set_capacity": (c_int, [c_void_p, ctypes.c_int64]),
    "phc_lb_clear_samples": (c_int, [c_void_p]),
    "phc_lb_sample_counts": (c_int, [c_void_p, i64p]),
    "phc_lb_get_samples": (c_int, [c_void_p, c_int, ctypes.c_int64, ctypes.c_int64, f32p]),
    "phc_lb_add_samples": (c_int, [c_void_p, c_int, f32p, ctypes.c_int64]),
    "phc_lb_collect": (c_int, [c_void_p, c_void_p, c_int, i32p, i32p, c_float, c_float, c_float,
                               ctypes.c_uint64]),
    "phc_lb_collect_onpolicy": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, i32p, c_float, c_float,
                                        c_float, ctypes.c_uint64]),
    "phc_lb_eval": (c_int, [c_void_p, ctypes.POINTER(c_float), c_void_p]),
    "phc_lb_train": (c_int, [c_void_p, c_int, c_int, c_float, ctypes.c_uint64, c_void_p]),
    "phc_lb_train_noalloc": (c_int, [c_void_p, c_int, c_int, c_float, ctypes.c_uint64, c_void_p]),
    "phc_lb_set_train_noise": (c_int, [c_void_p, c_float]),
    "phc_lb_infer": (c_int, [c_void_p, c_int, f32p, ctypes.c_int64, f32p]),
    "phc_fd_set_learned_boundary": (c_int, [c_void_p, c_void_p, c_float, c_float]),
    "phc_fd_lb_corrections": (c_int, [c_void_p, c_int, f32p]),
})


class LearnedBoundary:
    """Per-face MLP of the thin learned absorbing layer (phc_lb_*): parameters,
    training-sample store, GPU training and inference."""

    def __init__(self, seed=1):
        h = c_void_p()
        check(lib().phc_lb_create(seed, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(lib().phc_lb_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(LB_FACES * LB_NP, np.float32)
        check(lib().phc_lb_get_params(self.h, out))
        return out.reshape(LB_FACES, LB_NP)

    def set_params(self, theta):
        check(lib().phc_lb_set_params(self.h, np.ascontiguousarray(theta, np.float32).reshape(-1)))

    def set_capacity(self, per_face):
        check(lib().phc_lb_set_capacity(self.h, int(per_face)))

    def clear_samples(self):
        check(lib().phc_lb_clear_samples(self.h))

    def counts(self):
        c = np.zeros(LB_FACES, np.int64)
        check(lib().phc_lb_sample_counts(self.h, c))
        return c

    def samples(self, face):
        n = int(self.counts()[face])
        out = np.zeros(n * LB_SAMPLE, np.float32)
        if n:
            check(lib().phc_lb_get_samples(self.h, face, 0, n, out))
        return out.reshape(n, LB_SAMPLE)

    def add_samples(self, face, arr):
        a = np.ascontiguousarray(arr, np.float32).reshape(-1, LB_SAMPLE)
        check(lib().phc_lb_add_samples(self.h, face, a.reshape(-1), len(a)))

    def collect(self, wide, cur, origin, dims, halo=0.0, keep=1.0, s_min=0.0, seed=1):
        """Samples from a wide-sponge FdSolver's levels (cur = current level)."""
        check(lib().phc_lb_collect(self.h, wide.h, cur, np.asarray(origin, np.int32),
                                   np.asarray(dims, np.int32), halo, keep, s_min, seed))

    def collect_onpolicy(self, thin, wide, origin, halo=0.0, keep=1.0, s_min=0.0, seed=1):
        """Samples from a thin run at step n and a wide run at step n+1 (FdSolvers)."""
        check(lib().phc_lb_collect_onpolicy(self.h, thin.h, thin.i2, wide.h, wide.i2,
                                            np.asarray(origin, np.int32), halo, keep, s_min, seed))

    def eval(self, with_grad=True):
        loss = c_float()
        g = np.zeros(LB_FACES * LB_NP, np.float32) if with_grad else None
        check(lib().phc_lb_eval(self.h, ctypes.byref(loss), _ptr(g)))
        return loss.value, (g.reshape(LB_FACES, LB_NP) if with_grad else None)

    def set_train_noise(self, std):
        check(lib().phc_lb_set_train_noise(self.h, std))

    def train(self, steps, batch_per_face=4096, lr=1e-3, seed=1):
        hist = np.zeros(max(steps, 1), np.float32)
        check(lib().phc_lb_train(self.h, steps, batch_per_face, lr, seed, _ptr(hist)))
        return hist[:steps]

    def infer(self, face, samples):
        a = np.ascontiguousarray(samples, np.float32).reshape(-1, LB_SAMPLE)
        y = np.zeros(len(a), np.float32)
        check(lib().phc_lb_infer(self.h, face, a.reshape(-1), len(a), y))
        return y


def lb_corr_sizes(nx, ny, nz):
    n = (nx, ny, nz)
    sizes = []
    for f in range(LB_FACES):
        a = f >> 1
        nu = n[1] if a == 0 else n[0]
        nv = n[1] if a == 2 else n[2]
        sizes.append(LB_WIDTH * nu * nv)
    return sizes


# ----------------------------------------------------------------------- SEM
SEM_SEPARABLE = 0
SEM_ELEMENT = 1
SEM_ELEMENT_COLOR = 2
SEM_ELEMENT_MARCH = 3
SEM_ELEMENT_MFMA = 4


class SemParams(ctypes.Structure):
    _fields_ = [("ex", c_int), ("ey", c_int), ("ez", c_int),
                ("lx", c_float), ("ly", c_float), ("lz", c_float),
                ("order", c_int), ("model", c_void_p), ("dt", c_float),
                ("rhs", c_void_p), ("nsamples", c_int), ("src_element", c_int),
                ("kernel", c_int), ("weights_ref", c_int),
                ("tile_j", c_int), ("block_waves", c_int), ("chunk_i", c_int), ("index64", c_int),
                ("k_remainder", c_int)]


SIGNATURES.update({
    "phc_sem_create": (c_int, [ctypes.POINTER(SemParams), ctypes.POINTER(c_void_p)]),
    "phc_sem_destroy": (c_int, [c_void_p]),
    "phc_sem_sizes": (c_int, [c_void_p] + [ctypes.POINTER(c_int)] * 4),
    "phc_sem_compute_one_step": (c_int, [c_void_p, c_int, c_int, c_int]),
    "phc_sem_compute_one_step_rhs": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "phc_sem_set_source": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "phc_sem_run": (c_int, [c_void_p, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "phc_sem_set_receivers": (c_int, [c_void_p, c_int, c_void_p]),
    "phc_sem_get_traces": (c_int, [c_void_p, c_void_p, c_void_p]),
    "phc_sem_get_field": (c_int, [c_void_p, c_int, f32p]),
    "phc_sem_set_field": (c_int, [c_void_p, c_int, f32p]),
    "phc_sem_get_mass": (c_int, [c_void_p, f32p]),
    "phc_sem_apply_stiffness": (c_int, [c_void_p, f32p, f32p]),
    "phc_sem_check_finite": (c_int, [c_void_p, c_int]),
    "phc_sem_set_timing": (c_int, [c_void_p, c_int]),
    "phc_sem_last_run_kernel_ms": (c_int, [c_void_p, ctypes.POINTER(c_double)]),
    "phc_sem_synchronize": (c_int, [c_void_p]),
    "phc_sem_schedule_table": (c_int, [c_int, c_int, c_int, c_int, c_void_p, ctypes.POINTER(c_int)]),
    "phc_sem_autotune": (c_int, [c_void_p, c_int, ctypes.c_char_p, ctypes.POINTER(c_int),
                                 ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "phc_sem_element_from_point": (c_int, [c_int, c_int, c_int, c_float, c_float, c_float,
                                           c_float, c_float, c_float]),
})


SIGNATURES.update({
    "phc_sem2d_create": (c_int, [ctypes.POINTER(SemParams), ctypes.POINTER(c_void_p)]),
    "phc_sem2d_destroy": (c_int, [c_void_p]),
    "phc_sem2d_sizes": (c_int, [c_void_p] + [ctypes.POINTER(c_int)] * 3),
    "phc_sem2d_compute_one_step": (c_int, [c_void_p, c_int, c_int, c_int]),
    "phc_sem2d_compute_one_step_rhs": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "phc_sem2d_set_source": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "phc_sem2d_run": (c_int, [c_void_p, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "phc_sem2d_set_receivers": (c_int, [c_void_p, c_int, c_void_p]),
    "phc_sem2d_get_traces": (c_int, [c_void_p, c_void_p, c_void_p]),
    "phc_sem2d_get_field": (c_int, [c_void_p, c_int, f32p]),
    "phc_sem2d_set_field": (c_int, [c_void_p, c_int, f32p]),
    "phc_sem2d_get_mass": (c_int, [c_void_p, c_void_p, c_void_p]),
    "phc_sem2d_check_finite": (c_int, [c_void_p, c_int]),
    "phc_sem2d_set_timing": (c_int, [c_void_p, c_int]),
    "phc_sem2d_last_run_kernel_ms": (c_int, [c_void_p, ctypes.POINTER(c_double)]),
    "phc_sem2d_synchronize": (c_int, [c_void_p]),
    "phc_sem2d_autotune": (c_int, [c_void_p, c_int, ctypes.POINTER(c_int)]),
})


class Sem2dSolver:
    """SEMsolver of the reference's SEM2D build on the device: ex x ez
    quadrilaterals of order p on the x-z plane (SEMmesh with ey = 0), per-element
    wave speed, Ricker source at the first node of the element holding
    `src_point` (or `src_element`), paraxial boundary on the four edges."""

    def __init__(self, ex=65, lx=2000.0, order=3, dt=0.001, f0=10.0, model=None, nsamples=1000,
                 src_point=(1001.0, 1001.0, 1001.0), ez=None, lz=None, src_element=None,
                 weights_ref=True, source_order=1, rhs=None, chunk=0, index64=False):
        ez = ex if ez is None else ez
        lz = lx if lz is None else lz
        nE = ex * ez
        self.ex, self.ez, self.p = ex, ez, order
        self.model = (np.full(nE, 1500.0, np.float32) if model is None
                      else np.ascontiguousarray(model, np.float32))
        self.dt = np.float32(dt)
        self.rhs = (source_term(nsamples, self.dt, f0, source_order, EXP_FLOAT) if rhs is None
                    else np.ascontiguousarray(rhs, np.float32))
        if src_element is None:
            src_element = lib().phc_sem_element_from_point(ex, 0, ez, lx, 0.0, lz, *src_point)
        self.src_element = int(src_element)
        p = SemParams()
        p.ex, p.ey, p.ez = ex, 0, ez
        p.lx, p.ly, p.lz = lx, 0.0, lz
        p.order = order
        p.model = _ptr(self.model)
        p.dt = self.dt
        p.rhs = _ptr(self.rhs)
        p.nsamples = len(self.rhs)
        p.src_element = self.src_element
        p.kernel = SEM_SEPARABLE
        p.weights_ref = 1 if weights_ref else 0
        p.chunk_i = chunk
        p.index64 = 1 if index64 else 0
        h = c_void_p()
        check(lib().phc_sem2d_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        nn, nx, nz = c_int(), c_int(), c_int()
        check(lib().phc_sem2d_sizes(h, ctypes.byref(nn), ctypes.byref(nx), ctypes.byref(nz)))
        self.nn, self.nx, self.nz = nn.value, nx.value, nz.value
        exi, ezi = self.src_element % ex, self.src_element // ex
        self.src_node = exi * order + ezi * order * self.nx
        self.i1, self.i2 = 0, 1
        self.nrec = 0
        self.it = 0

    def close(self):
        if self.h:
            check(lib().phc_sem2d_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute_one_step(self, time_sample, i1, i2):
        check(lib().phc_sem2d_compute_one_step(self.h, time_sample, i1, i2))

    def step(self):
        """One iteration of SEMproxy::run (SEMproxy.cpp:45-55)."""
        self.compute_one_step(self.it, self.i1, self.i2)
        self.i1, self.i2 = self.i2, self.i1
        self.it += 1

    def run(self, nsteps):
        i1, i2 = c_int(self.i1), c_int(self.i2)
        check(lib().phc_sem2d_run(self.h, self.it, nsteps, ctypes.byref(i1), ctypes.byref(i2)))
        self.i1, self.i2 = i1.value, i2.value
        self.it += nsteps
        src = np.zeros(nsteps, np.float32)
        rec = np.zeros(max(self.nrec, 1) * nsteps, np.float32)
        check(lib().phc_sem2d_get_traces(self.h, _ptr(src), _ptr(rec) if self.nrec else None))
        return src, (rec.reshape(self.nrec, nsteps) if self.nrec else None)

    def set_receivers(self, nodes):
        a = np.ascontiguousarray(nodes, np.int32)
        self.nrec = len(a)
        check(lib().phc_sem2d_set_receivers(self.h, self.nrec, _ptr(a)))

    def field(self, level):
        out = np.zeros(self.nn, np.float32)
        check(lib().phc_sem2d_get_field(self.h, level, out))
        return out

    def set_field(self, level, arr):
        check(lib().phc_sem2d_set_field(self.h, level, np.ascontiguousarray(arr, np.float32)))

    def current(self):
        return self.field(self.i2)

    def previous(self):
        return self.field(self.i1)

    def mass(self):
        """(lumped mass M, boundary term S) per node."""
        m = np.zeros(self.nn, np.float32)
        b = np.zeros(self.nn, np.float32)
        check(lib().phc_sem2d_get_mass(self.h, _ptr(m), _ptr(b)))
        return m, b

    def set_timing(self, on):
        check(lib().phc_sem2d_set_timing(self.h, 1 if on else 0))

    def last_kernel_ms(self):
        v = c_double()
        check(lib().phc_sem2d_last_run_kernel_ms(self.h, ctypes.byref(v)))
        return v.value

    def synchronize(self):
        check(lib().phc_sem2d_synchronize(self.h))

    def check_finite(self, level):
        return lib().phc_sem2d_check_finite(self.h, level)

    def autotune(self, steps_per_trial=10):
        c = c_int()
        check(lib().phc_sem2d_autotune(self.h, steps_per_trial, ctypes.byref(c)))
        return c.value


SIGNATURES.update({
    "phc_sem_color_count": (c_int, [c_int, c_int, c_int, ctypes.POINTER(c_int)]),
    "phc_sem_color_elements": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p]),
})


def sem_color_elements(ex, ey, ez):
    """Element lists per colour (SEMmesh::sortElementsByColor; 8 colours in 3D,
    the reference's 4 when ey == 0)."""
    mpc = c_int()
    nc = lib().phc_sem_color_count(ex, ey, ez, ctypes.byref(mpc))
    if nc == 0:
        raise PhcError("bad mesh")
    counts = np.zeros(nc, np.int32)
    lists = np.zeros(nc * mpc.value, np.int32)
    check(lib().phc_sem_color_elements(ex, ey, ez, _ptr(counts), _ptr(lists)))
    return [lists[c * mpc.value:c * mpc.value + counts[c]] for c in range(nc)]


def sem_model_gaussian(ex, ey, ez, sigma=5.0, vmin=2000.0, vmax=4000.0, seed=1):
    out = np.zeros(ex * ey * ez, np.float32)
    check(lib().phc_sem_model_gaussian(ex, ey, ez, sigma, vmin, vmax, seed, out))
    return out


SIGNATURES.update({
    "phc_sem64_create": (c_int, [ctypes.POINTER(SemParams), ctypes.POINTER(c_void_p)]),
    "phc_sem64_destroy": (c_int, [c_void_p]),
    "phc_sem64_set_receivers": (c_int, [c_void_p, c_int, c_void_p]),
    "phc_sem64_run": (c_int, [c_void_p, c_int, c_int]),
    "phc_sem64_get_traces": (c_int, [c_void_p, c_void_p, c_void_p]),
    "phc_sem64_get_field": (c_int, [c_void_p, c_int, f64p]),
})


class Sem64Solver:
    """The SEM loop in FP64 (accuracy yardstick), same arguments as SemSolver."""

    def __init__(self, ex=10, lx=2000.0, order=2, dt=0.001, f0=10.0, model=None, nsamples=100,
                 src_point=(1001.0, 1001.0, 1001.0), ey=None, ez=None, ly=None, lz=None,
                 weights_ref=False, source_order=1, rhs=None):
        ey = ex if ey is None else ey
        ez = ex if ez is None else ez
        ly = lx if ly is None else ly
        lz = lx if lz is None else lz
        nE = ex * ey * ez
        self.model = (np.full(nE, 1500.0, np.float32) if model is None
                      else np.ascontiguousarray(model, np.float32))
        self.dt = np.float32(dt)
        self.rhs = (source_term(nsamples, self.dt, f0, source_order, EXP_FLOAT) if rhs is None
                    else np.ascontiguousarray(rhs, np.float32))
        self.src_element = lib().phc_sem_element_from_point(ex, ey, ez, lx, ly, lz, *src_point)
        p = SemParams()
        p.ex, p.ey, p.ez = ex, ey, ez
        p.lx, p.ly, p.lz = lx, ly, lz
        p.order = order
        p.model = _ptr(self.model)
        p.dt = self.dt
        p.rhs = _ptr(self.rhs)
        p.nsamples = len(self.rhs)
        p.src_element = self.src_element
        p.weights_ref = 1 if weights_ref else 0
        h = c_void_p()
        check(lib().phc_sem64_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.nx, self.ny, self.nz = ex * order + 1, ey * order + 1, ez * order + 1
        self.nn = self.nx * self.ny * self.nz
        self.nrec = 0
        self.it = 0

    def close(self):
        if self.h:
            check(lib().phc_sem64_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_receivers(self, nodes):
        a = np.ascontiguousarray(nodes, np.int32)
        self.nrec = len(a)
        check(lib().phc_sem64_set_receivers(self.h, self.nrec, _ptr(a)))

    def run(self, nsteps):
        check(lib().phc_sem64_run(self.h, self.it, nsteps))
        self.it += nsteps
        src = np.zeros(nsteps)
        rec = np.zeros(max(self.nrec, 1) * nsteps)
        check(lib().phc_sem64_get_traces(self.h, _ptr(src), _ptr(rec) if self.nrec else None))
        return src, (rec.reshape(self.nrec, nsteps) if self.nrec else None)

    def field(self, current=True):
        out = np.zeros(self.nn)
        check(lib().phc_sem64_get_field(self.h, 1 if current else 0, out))
        return out


def sem_schedule_table(march_planes, tile_columns, cus=256, split=1):
    """The per-XCD schedule table of the separable SEM kernel (host only):
    rows (tile column, first plane, end plane), row b read by hardware block b."""
    n = c_int(0)
    check(lib().phc_sem_schedule_table(march_planes, tile_columns, cus, split, None, ctypes.byref(n)))
    out = np.zeros((n.value, 3), np.int32)
    check(lib().phc_sem_schedule_table(march_planes, tile_columns, cus, split, _ptr(out), ctypes.byref(n)))
    return out


class SemSolver:
    """SEMsolver (src/sem/solver/SEMsolver.hpp:18-91) on the device: the
    structured box of SEMmesh, per-element wave speed, Ricker source in the
    element containing `src_point`, levels i1/i2 as in SEMproxy."""

    def __init__(self, ex=10, lx=2000.0, order=2, dt=0.001, f0=10.0, model=None, nsamples=100,
                 src_point=(1001.0, 1001.0, 1001.0), ey=None, ez=None, ly=None, lz=None,
                 kernel=SEM_SEPARABLE, weights_ref=True, source_order=1, rhs=None,
                 tile_j=0, block_waves=0, chunk_i=0, index64=False, k_remainder=0):
        ey = ex if ey is None else ey
        ez = ex if ez is None else ez
        ly = lx if ly is None else ly
        lz = lx if lz is None else lz
        nE = ex * ey * ez
        self.model = (np.full(nE, 1500.0, np.float32) if model is None
                      else np.ascontiguousarray(model, np.float32))
        self.dt = np.float32(dt)
        self.rhs = (source_term(nsamples, self.dt, f0, source_order, EXP_FLOAT) if rhs is None
                    else np.ascontiguousarray(rhs, np.float32))
        self.src_element = lib().phc_sem_element_from_point(ex, ey, ez, lx, ly, lz, *src_point)
        p = SemParams()
        p.ex, p.ey, p.ez = ex, ey, ez
        p.lx, p.ly, p.lz = lx, ly, lz
        p.order = order
        p.model = _ptr(self.model)
        p.dt = self.dt
        p.rhs = _ptr(self.rhs)
        p.nsamples = len(self.rhs)
        p.src_element = self.src_element
        p.kernel = kernel
        p.weights_ref = 1 if weights_ref else 0
        p.tile_j, p.block_waves, p.chunk_i = tile_j, block_waves, chunk_i
        p.index64 = 1 if index64 else 0
        p.k_remainder = k_remainder
        h = c_void_p()
        check(lib().phc_sem_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        nn, nx, ny, nz = c_int(), c_int(), c_int(), c_int()
        check(lib().phc_sem_sizes(h, ctypes.byref(nn), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
        self.nn, self.nx, self.ny, self.nz = nn.value, nx.value, ny.value, nz.value
        self.i1, self.i2 = 0, 1
        self.nrec = 0
        self.it = 0

    def close(self):
        if self.h:
            check(lib().phc_sem_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute_one_step(self, time_sample, i1, i2):
        check(lib().phc_sem_compute_one_step(self.h, time_sample, i1, i2))

    def step(self):
        """One iteration of SEMproxy::run (SEMproxy.cpp:45-55)."""
        self.compute_one_step(self.it, self.i1, self.i2)
        self.i1, self.i2 = self.i2, self.i1
        self.it += 1

    def run(self, nsteps):
        i1, i2 = c_int(self.i1), c_int(self.i2)
        check(lib().phc_sem_run(self.h, self.it, nsteps, ctypes.byref(i1), ctypes.byref(i2)))
        self.i1, self.i2 = i1.value, i2.value
        self.it += nsteps
        src = np.zeros(nsteps, np.float32)
        rec = np.zeros(max(self.nrec, 1) * nsteps, np.float32)
        check(lib().phc_sem_get_traces(self.h, _ptr(src), _ptr(rec) if self.nrec else None))
        return src, (rec.reshape(self.nrec, nsteps) if self.nrec else None)

    def set_receivers(self, nodes):
        a = np.ascontiguousarray(nodes, np.int32)
        self.nrec = len(a)
        check(lib().phc_sem_set_receivers(self.h, self.nrec, _ptr(a)))

    def field(self, level):
        out = np.zeros(self.nn, np.float32)
        check(lib().phc_sem_get_field(self.h, level, out))
        return out

    def set_field(self, level, arr):
        check(lib().phc_sem_set_field(self.h, level, np.ascontiguousarray(arr, np.float32)))

    def current(self):
        return self.field(self.i2)

    def previous(self):
        return self.field(self.i1)

    def mass(self):
        out = np.zeros(self.nn, np.float32)
        check(lib().phc_sem_get_mass(self.h, out))
        return out

    def apply_stiffness(self, p):
        y = np.zeros(self.nn, np.float32)
        check(lib().phc_sem_apply_stiffness(self.h, np.ascontiguousarray(p, np.float32), y))
        return y

    def set_timing(self, on):
        check(lib().phc_sem_set_timing(self.h, 1 if on else 0))

    def last_kernel_ms(self):
        v = c_double()
        check(lib().phc_sem_last_run_kernel_ms(self.h, ctypes.byref(v)))
        return v.value

    def synchronize(self):
        check(lib().phc_sem_synchronize(self.h))

    def check_finite(self, level):
        return lib().phc_sem_check_finite(self.h, level)

    def autotune(self, steps_per_trial=10, cache_path=None):
        tj, ty, ci = c_int(), c_int(), c_int()
        cp = None if cache_path is None else cache_path.encode()
        check(lib().phc_sem_autotune(self.h, steps_per_trial, cp, ctypes.byref(tj), ctypes.byref(ty),
                                     ctypes.byref(ci)))
        return tj.value, ty.value, ci.value
