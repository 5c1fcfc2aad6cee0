"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  The C sources beside this
file restate the reference's hot path; see oracle.h for the file:line map.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class FdGrid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "nx", "ny", "nz", "lx", "ly", "lz", "xs", "ys", "zs",
        "x3", "x4", "y3", "y4", "z3", "z4")]


class SemMesh(ctypes.Structure):
    _fields_ = [("ex", ctypes.c_int), ("ey", ctypes.c_int), ("ez", ctypes.c_int),
                ("lx", ctypes.c_float), ("ly", ctypes.c_float), ("lz", ctypes.c_float),
                ("order", ctypes.c_int),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("hx", ctypes.c_float), ("hy", ctypes.c_float), ("hz", ctypes.c_float)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_float, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_double
        P = ctypes.POINTER
        L.orc_ricker.restype = c_float
        L.orc_ricker.argtypes = [c_float, c_float, c_int, c_int]
        L.orc_source_term.restype = None
        L.orc_source_term.argtypes = [c_int, c_float, c_float, c_int, c_int, f32p]
        L.orc_ricker_f64.restype = c_double
        L.orc_ricker_f64.argtypes = [c_double, c_double, c_int]
        L.orc_fd_init_coef.restype = None
        L.orc_fd_init_coef.argtypes = [c_int, c_float, f32p]
        L.orc_fd_coef0.restype = c_double
        L.orc_fd_coef0.argtypes = [f32p, c_int, f32p, c_int, f32p, c_int]
        L.orc_fd_compute_dt.restype = c_float
        L.orc_fd_compute_dt.argtypes = [c_float, f32p, c_int, f32p, c_int, f32p, c_int, c_float]
        L.orc_fd_sponge.restype = None
        L.orc_fd_sponge.argtypes = [c_int, c_int, c_int, c_int, c_float, f32p]
        L.orc_fd_sponge_table.restype = None
        L.orc_fd_sponge_table.argtypes = [c_int, c_float, f32p]
        L.orc_fd_add_rhs.restype = None
        L.orc_fd_add_rhs.argtypes = [P(FdGrid), f32p, f32p, c_float]
        L.orc_fd_inner3d.restype = None
        L.orc_fd_inner3d.argtypes = [P(FdGrid), f32p, f32p, f32p, c_double, f32p, f32p, f32p]
        L.orc_fd_apply_sponge.restype = None
        L.orc_fd_apply_sponge.argtypes = [P(FdGrid), f32p, f32p, f32p]
        L.orc_fd_run.restype = None
        L.orc_fd_run.argtypes = [P(FdGrid), f32p, f32p, f32p, c_double, f32p, f32p, f32p,
                                 c_int, c_int, f32p, f32p, P(c_int), ctypes.c_void_p,
                                 c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_fd_pml_profile.restype = None
        L.orc_fd_pml_profile.argtypes = [c_int, c_int, c_float, f32p]
        L.orc_fd_pml_scale.restype = c_float
        L.orc_fd_pml_scale.argtypes = [c_float, c_float, c_int, c_float]
        L.orc_fd_run_pml.restype = None
        L.orc_fd_run_pml.argtypes = [P(FdGrid), i32p, f32p, f32p, f32p, c_double, f32p, f32p, f32p, f32p,
                                     f32p, f32p, f32p, c_int, c_int, f32p, f32p, P(c_int), ctypes.c_void_p,
                                     c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_fd_run_f64.restype = None
        L.orc_fd_run_f64.argtypes = [P(FdGrid), f64p, f64p, f64p, c_double, f64p, f64p, f64p,
                                     c_int, c_int, f64p, f64p, P(c_int), ctypes.c_void_p,
                                     c_int, ctypes.c_void_p, ctypes.c_void_p]
        _bind_sem(L)
        _lib = L
    return _lib


def _bind_sem(L):
    c_int, c_float, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_double
    P = ctypes.POINTER
    if not hasattr(L, "orc_gll_points"):
        return
    L.orc_gll_points.restype = None
    L.orc_gll_points.argtypes = [c_int, f64p]
    L.orc_gll_weights.restype = None
    L.orc_gll_weights.argtypes = [c_int, f64p, c_int]
    L.orc_gll_deriv.restype = None
    L.orc_gll_deriv.argtypes = [c_int, f64p, f64p]
    L.orc_sem_mesh_init.restype = None
    L.orc_sem_mesh_init.argtypes = [P(SemMesh), c_int, c_int, c_int, c_float, c_float, c_float, c_int]
    L.orc_sem_coord_1d.restype = None
    L.orc_sem_coord_1d.argtypes = [c_int, c_int, c_int, f32p]
    L.orc_sem_nodes_list.restype = None
    L.orc_sem_nodes_list.argtypes = [P(SemMesh), i32p]
    L.orc_sem_element_from_point.restype = c_int
    L.orc_sem_element_from_point.argtypes = [P(SemMesh), c_float, c_float, c_float]
    L.orc_sem_interior_nodes.restype = c_int
    L.orc_sem_interior_nodes.argtypes = [P(SemMesh), i32p]
    L.orc_sem_create.restype = ctypes.c_void_p
    L.orc_sem_create.argtypes = [P(SemMesh), f32p, c_float]
    L.orc_sem_destroy.restype = None
    L.orc_sem_destroy.argtypes = [ctypes.c_void_p]
    L.orc_sem_num_nodes.restype = c_int
    L.orc_sem_num_nodes.argtypes = [ctypes.c_void_p]
    L.orc_sem_step.restype = None
    L.orc_sem_step.argtypes = [ctypes.c_void_p, f32p, f32p, c_int, c_float]
    L.orc_sem_get_my.restype = None
    L.orc_sem_get_my.argtypes = [ctypes.c_void_p, f32p, f32p]
    L.orc_sem_set_threads.restype = None
    L.orc_sem_set_threads.argtypes = [ctypes.c_void_p, c_int]
    L.orc_sem_apply_k.restype = None
    L.orc_sem_apply_k.argtypes = [ctypes.c_void_p, f32p, f32p, f32p]
    L.orc_sem64_create.restype = ctypes.c_void_p
    L.orc_sem64_create.argtypes = [P(SemMesh), f64p, c_double]
    L.orc_sem64_destroy.restype = None
    L.orc_sem64_destroy.argtypes = [ctypes.c_void_p]
    L.orc_sem64_step.restype = None
    L.orc_sem64_step.argtypes = [ctypes.c_void_p, f64p, f64p, c_int, c_double]
    L.orc_sem64_apply_k.restype = None
    L.orc_sem64_apply_k.argtypes = [ctypes.c_void_p, f64p, f64p, f64p]
    L.orc_sem_dense_element_f64.restype = None
    L.orc_sem_dense_element_f64.argtypes = [c_int, f64p, f64p, f64p, f64p, f64p]
    V = ctypes.c_void_p
    L.orc_sem2d_create.restype = V
    L.orc_sem2d_create.argtypes = [c_int, c_int, c_float, c_float, c_int, f32p, c_float]
    L.orc_sem2d_destroy.restype = None
    L.orc_sem2d_destroy.argtypes = [V]
    L.orc_sem2d_sizes.restype = None
    L.orc_sem2d_sizes.argtypes = [V, i32p]
    L.orc_sem2d_lists.restype = None
    L.orc_sem2d_lists.argtypes = [V, V, V, V, V, V]
    L.orc_sem2d_coords.restype = None
    L.orc_sem2d_coords.argtypes = [V, f32p, f32p]
    L.orc_sem2d_step.restype = None
    L.orc_sem2d_step.argtypes = [V, f32p, f32p, c_int, c_float]
    L.orc_sem2d_apply_k.restype = None
    L.orc_sem2d_apply_k.argtypes = [V, f32p, f32p, f32p]
    L.orc_sem2d_get_my.restype = None
    L.orc_sem2d_get_my.argtypes = [V, V, V, V]
    L.orc_sem2d_boundary_term.restype = None
    L.orc_sem2d_boundary_term.argtypes = [V, f32p]
    L.orc_sem2d_set_paraxial.restype = None
    L.orc_sem2d_set_paraxial.argtypes = [V, c_int]
    L.orc_sem2d_colors.restype = c_int
    L.orc_sem2d_colors.argtypes = [c_int, c_int, i32p, i32p]
    L.orc_sem2d64_create.restype = V
    L.orc_sem2d64_create.argtypes = [c_int, c_int, c_float, c_float, c_int, f64p, c_double, c_int]
    L.orc_sem2d64_destroy.restype = None
    L.orc_sem2d64_destroy.argtypes = [V]
    L.orc_sem2d64_step.restype = None
    L.orc_sem2d64_step.argtypes = [V, f64p, f64p, c_double]


def _vp(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------------ FD
def fd_coefs(L, dx):
    c = np.zeros(L + 1, np.float32)
    lib().orc_fd_init_coef(L, dx, c)
    return c


class FdProblem:
    """Everything the reference's FDTDInit builds (FDTDinit.hpp:36-396),
    restated.  `model` is "layered2" (the reference's init_models: vmin^2 dt^2
    above nz/2, twice that below), "homogeneous" (vmin everywhere, the model of
    the published test_20240512 logs) or an explicit float32 array of vp^2 dt^2
    in IDX3 order."""

    def __init__(self, nx=150, ny=None, nz=None, L=4, dx=10.0, xs=None, ys=None, zs=None,
                 f0=15.0, time_max=1.0, vmin=1500.0, vmax=4500.0, cfl=0.8,
                 source_order=1, sponge_width=20, sponge_alpha=-0.00015,
                 model="layered2", dt=None, nsamples=None, pml=False, ndamp=None, ly=None, lz=None):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self.nx, self.ny, self.nz, self.L = nx, ny, nz, L
        # per-axis half widths (FDTDinit.hpp:96-110); L is lx
        ly = L if ly is None else ly
        lz = L if lz is None else lz
        self.lh = (L, ly, lz)
        xs = nx // 2 if xs is None else xs
        ys = ny // 2 if ys is None else ys
        zs = nz // 2 if zs is None else zs
        g = FdGrid()
        g.nx, g.ny, g.nz = nx, ny, nz
        g.lx, g.ly, g.lz = L, ly, lz
        g.xs, g.ys, g.zs = xs, ys, zs
        g.x3, g.x4, g.y3, g.y4, g.z3, g.z4 = 0, nx, 0, ny, 0, nz
        self.g = g
        self.cx = fd_coefs(L, dx)
        self.cy = fd_coefs(ly, dx)
        self.cz = fd_coefs(lz, dx)
        self.coef0 = lib().orc_fd_coef0(self.cx, L, self.cy, ly, self.cz, lz)
        if dt is None:
            dt = lib().orc_fd_compute_dt(vmax, self.cx, L, self.cy, ly, self.cz, lz, cfl)
        self.dt = np.float32(dt)
        if nsamples is None:
            nsamples = int(np.float32(time_max) / self.dt)
        self.nsamples = nsamples
        self.f0 = f0
        self.rhs = np.zeros(nsamples, np.float32)
        lib().orc_source_term(nsamples, self.dt, f0, source_order, 1, self.rhs)
        v0 = np.float32(np.float32(np.float32(vmin) * np.float32(vmin)) * self.dt) * self.dt
        if isinstance(model, np.ndarray):
            self.vp = np.ascontiguousarray(model, dtype=np.float32).reshape(-1)
        elif model == "homogeneous":
            self.vp = np.full(nx * ny * nz, v0, np.float32)
        elif model == "layered2":
            vp = np.full((nx, ny, nz), v0, np.float32)
            vp[:, :, nz // 2:] = np.float32(2) * v0
            self.vp = vp.reshape(-1)
        else:
            raise ValueError(model)
        self.sponge = np.zeros(nx * ny * nz, np.float32)
        lib().orc_fd_sponge(nx, ny, nz, sponge_width, sponge_alpha, self.sponge)
        self.sponge_width = sponge_width
        self.sponge_alpha = sponge_alpha
        # PML (FDTDinit.hpp:155-198, FDTDutils.hpp:129-171): width ntaper *
        # lambdamax / dx with ntaper = 4, lambdamax = vmin / (2.5 f0)
        self.pml = pml
        if pml:
            fmax = np.float32(np.float32(2.5) * np.float32(f0))
            lam = np.float32(np.float32(vmin) / fmax)
            nd = int(np.float32(np.float32(4 * lam) / np.float32(dx))) if ndamp is None else ndamp
            self.ndamp = np.array([nd, nd, nd], np.int32)
            self.eta = []
            for n in (nx, ny, nz):
                prof = np.zeros(n + 2, np.float32)
                if nd > 0:
                    lib().orc_fd_pml_profile(n, nd, lib().orc_fd_pml_scale(self.dt, vmax, nd, dx), prof)
                self.eta.append(prof)
            self.hd2 = np.full(3, np.float32(1.0 / (4.0 * dx * dx)), np.float32)
            self.phi = np.zeros(nx * ny * nz, np.float32)
        ext = (nx + 2 * L) * (ny + 2 * ly) * (nz + 2 * lz)
        self.p0 = np.full(ext, np.float32(0.000001), np.float32)
        self.p1 = np.full(ext, np.float32(0.000001), np.float32)
        self.cur_is_p1 = ctypes.c_int(1)
        self.it = 0

    def ext_shape(self):
        lx, ly, lz = self.lh
        return (self.nx + 2 * lx, self.ny + 2 * ly, self.nz + 2 * lz)

    def run(self, nsteps, receivers=None):
        """Run nsteps reference steps; returns (source trace, receiver traces)."""
        tr = np.zeros(nsteps, np.float32)
        nrec = 0 if receivers is None else len(receivers)
        rec = None if not nrec else np.ascontiguousarray(receivers, np.int32).reshape(-1)
        rv = None if not nrec else np.zeros(nrec * nsteps, np.float32)
        if self.pml:
            lib().orc_fd_run_pml(ctypes.byref(self.g), self.ndamp, self.cx, self.cy, self.cz, self.coef0,
                                 self.vp, self.eta[0], self.eta[1], self.eta[2], self.hd2, self.phi, self.rhs,
                                 self.it, nsteps, self.p0, self.p1, ctypes.byref(self.cur_is_p1), _vp(tr), nrec,
                                 _vp(rec), _vp(rv))
        else:
            lib().orc_fd_run(ctypes.byref(self.g), self.cx, self.cy, self.cz, self.coef0, self.vp,
                             self.sponge, self.rhs, self.it, nsteps, self.p0, self.p1,
                             ctypes.byref(self.cur_is_p1), _vp(tr), nrec, _vp(rec), _vp(rv))
        self.it += nsteps
        return tr, (None if rv is None else rv.reshape(nrec, nsteps))

    def current(self):
        return (self.p1 if self.cur_is_p1.value else self.p0).reshape(self.ext_shape())

    def previous(self):
        return (self.p0 if self.cur_is_p1.value else self.p1).reshape(self.ext_shape())


def fd_exact_coefs(L, dx):
    """Exact central second-difference weights (double) / dx^2, L <= 5."""
    from fractions import Fraction as F
    tab = {1: [F(-2), F(1)],
           2: [F(-5, 2), F(4, 3), F(-1, 12)],
           3: [F(-49, 18), F(3, 2), F(-3, 20), F(1, 90)],
           4: [F(-205, 72), F(8, 5), F(-1, 5), F(8, 315), F(-1, 560)],
           5: [F(-5269, 1800), F(5, 3), F(-5, 21), F(5, 126), F(-5, 1008), F(1, 3150)]}
    return np.array([float(c) for c in tab[L]], np.float64) / (dx * dx)


# ----------------------------------------------------------------------- SEM
def gll(p, ref_rounding=True):
    """(points, weights, dphi[j, i] = l_j'(xi_i)) of order p."""
    xi = np.zeros(p + 1)
    w = np.zeros(p + 1)
    d = np.zeros((p + 1) * (p + 1))
    lib().orc_gll_points(p, xi)
    lib().orc_gll_weights(p, w, 1 if ref_rounding else 0)
    lib().orc_gll_deriv(p, xi, d)
    return xi, w, d.reshape(p + 1, p + 1)


class SemProblem:
    """The reference SEM proxy (SEMproxy.cpp:10-119, SEMsolver.cpp:21-212)
    restated: structured box mesh of ex^3 elements of order p, per-element
    velocity, Ricker source in the element containing `src_point`, leapfrog
    with Dirichlet boundary nodes.  `fp64=True` uses the FP64 oracle."""

    def __init__(self, ex=10, lx=2000.0, order=2, dt=0.001, f0=10.0, model=None, nsamples=100,
                 src_point=(1001.0, 1001.0, 1001.0), ey=None, ez=None, ly=None, lz=None,
                 fp64=False, source_order=1):
        ey = ex if ey is None else ey
        ez = ex if ez is None else ez
        ly = lx if ly is None else ly
        lz = lx if lz is None else lz
        m = SemMesh()
        lib().orc_sem_mesh_init(ctypes.byref(m), ex, ey, ez, lx, ly, lz, order)
        self.mesh = m
        self.p = order
        self.nE = ex * ey * ez
        self.nN = m.nx * m.ny * m.nz
        self.dt = np.float32(dt)
        if model is None:
            model = np.full(self.nE, 1500.0, np.float32)
        self.model = np.ascontiguousarray(model, np.float32)
        self.src_element = lib().orc_sem_element_from_point(ctypes.byref(m), *src_point)
        self.rhs = np.zeros(nsamples, np.float32)
        # SEM translation unit: std::exp(float)
        lib().orc_source_term(nsamples, self.dt, f0, source_order, 0, self.rhs)
        self.fp64 = fp64
        if fp64:
            self.h = lib().orc_sem64_create(ctypes.byref(m), self.model.astype(np.float64), float(self.dt))
            self.p0 = np.zeros(self.nN)
            self.p1 = np.zeros(self.nN)
        else:
            self.h = lib().orc_sem_create(ctypes.byref(m), self.model, self.dt)
            self.p0 = np.zeros(self.nN, np.float32)
            self.p1 = np.zeros(self.nN, np.float32)
        self.nodes = np.zeros(self.nE * (order + 1) ** 3, np.int32)
        lib().orc_sem_nodes_list(ctypes.byref(m), self.nodes)
        self.nodes = self.nodes.reshape(self.nE, (order + 1) ** 3)
        self.src_node = int(self.nodes[self.src_element, 0])
        self.i1, self.i2 = 0, 1
        self.it = 0

    def set_threads(self, n):
        """n > 1: the element loop colour by colour on n OpenMP threads (no
        race; FP32 sums in colour order, so the last bits can differ from
        the sequential loop: for timing)."""
        if self.fp64:
            raise ValueError("the FP64 oracle runs sequentially")
        lib().orc_sem_set_threads(self.h, int(n))

    def __del__(self):
        try:
            if self.fp64:
                lib().orc_sem64_destroy(self.h)
            else:
                lib().orc_sem_destroy(self.h)
        except Exception:
            pass

    def level(self, i):
        return self.p0 if i == 0 else self.p1

    def run(self, nsteps, receivers=None):
        """SEMproxy::run: computeOneStep(i1, i2) -> output(i1) -> swap.  Returns
        the source-node trace (value after each step, outputPnValues) and
        receiver traces (node ids)."""
        tr = np.zeros(nsteps, self.p0.dtype)
        rec = None if receivers is None else np.zeros((len(receivers), nsteps), self.p0.dtype)
        for s in range(nsteps):
            it = self.it + s
            a, b = self.level(self.i1), self.level(self.i2)
            if self.fp64:
                lib().orc_sem64_step(self.h, a, b, self.src_element, float(self.rhs[it]))
            else:
                lib().orc_sem_step(self.h, a, b, self.src_element, self.rhs[it])
            tr[s] = a[self.src_node]
            if rec is not None:
                rec[:, s] = a[np.asarray(receivers)]
            self.i1, self.i2 = self.i2, self.i1
        self.it += nsteps
        return tr, rec

    def current(self):
        return self.level(self.i2)

    def previous(self):
        return self.level(self.i1)

    def apply_k(self, p):
        if self.fp64:
            y = np.zeros(self.nN)
            M = np.zeros(self.nN)
            lib().orc_sem64_apply_k(self.h, np.ascontiguousarray(p, np.float64), y, M)
        else:
            y = np.zeros(self.nN, np.float32)
            M = np.zeros(self.nN, np.float32)
            lib().orc_sem_apply_k(self.h, np.ascontiguousarray(p, np.float32), y, M)
        return y, M

    def node_coords(self):
        """Physical coordinates of every node (x, y, z), node id = ix + iz nx + iy nx nz."""
        m = self.mesh
        xi = gll(self.p)[0]
        def axis(ne, h):
            c = np.zeros(ne * self.p + 1)
            for e in range(ne):
                c[e * self.p:(e + 1) * self.p + 1] = e * h + (xi + 1) * h / 2
            return c
        cx = axis(m.ex, m.lx / m.ex)
        cy = axis(m.ey, m.ly / m.ey)
        cz = axis(m.ez, m.lz / m.ez)
        n = np.arange(self.nN)
        ix = n % m.nx
        iz = (n // m.nx) % m.nz
        iy = n // (m.nx * m.nz)
        return cx[ix], cy[iy], cz[iz]


class Sem2dProblem:
    """The reference SEM proxy built with SEM2D (SEMproxy.cpp:10-119 with
    ey = 0, SEMsolver.cpp:21-125): ex x ez quadrilaterals of order p on the x-z
    plane, per-element velocity, Ricker source at the first node of
    `src_element` (or of the element holding `src_point`), paraxial boundary."""

    def __init__(self, ex=65, ez=None, lx=2000.0, lz=None, order=3, dt=0.001, f0=10.0, model=None,
                 nsamples=1000, src_element=None, src_point=(1001.0, 1001.0), source_order=1,
                 paraxial=True):
        ez = ex if ez is None else ez
        lz = lx if lz is None else lz
        self.ex, self.ez, self.p = ex, ez, order
        self.nE = ex * ez
        if model is None:
            model = np.full(self.nE, 1500.0, np.float32)
        self.model = np.ascontiguousarray(model, np.float32)
        self.dt = np.float32(dt)
        self.h = lib().orc_sem2d_create(ex, ez, lx, lz, order, self.model, self.dt)
        if not self.h:
            raise ValueError("bad 2D mesh")
        if not paraxial:
            lib().orc_sem2d_set_paraxial(self.h, 0)
        sz = np.zeros(7, np.int32)
        lib().orc_sem2d_sizes(self.h, sz)
        self.nx, self.nz, self.nN, _, self.nI, self.nB, self.nF = (int(v) for v in sz)
        q = order + 1
        self.nodes = np.zeros(self.nE * q * q, np.int32)
        self.interior = np.zeros(self.nI, np.int32)
        self.bnodes = np.zeros(self.nB, np.int32)
        self.faces = np.zeros(self.nF * (2 + q), np.int32)
        self.fmap = np.zeros(self.nF * q, np.int32)
        lib().orc_sem2d_lists(self.h, _vp(self.nodes), _vp(self.interior), _vp(self.bnodes),
                              _vp(self.faces), _vp(self.fmap))
        self.nodes = self.nodes.reshape(self.nE, q * q)
        self.faces = self.faces.reshape(self.nF, 2 + q)
        self.fmap = self.fmap.reshape(self.nF, q)
        if src_element is None:
            hx, hz = np.float32(lx / ex), np.float32(lz / ez)
            ix = min(int(np.float32(src_point[0]) // hx), ex - 1)
            iz = min(int(np.float32(src_point[1]) // hz), ez - 1)
            src_element = ix + iz * ex
        self.src_element = int(src_element)
        self.src_node = int(self.nodes[self.src_element, 0])
        self.rhs = np.zeros(nsamples, np.float32)
        lib().orc_source_term(nsamples, self.dt, f0, source_order, 0, self.rhs)
        self.p0 = np.zeros(self.nN, np.float32)
        self.p1 = np.zeros(self.nN, np.float32)
        self.i1, self.i2 = 0, 1
        self.it = 0

    def __del__(self):
        try:
            lib().orc_sem2d_destroy(self.h)
        except Exception:
            pass

    def level(self, i):
        return self.p0 if i == 0 else self.p1

    def run(self, nsteps, receivers=None):
        """computeOneStep(i1, i2) -> output(i1) -> swap, as SEMproxy::run does;
        returns the source-node trace and the receiver traces (node ids)."""
        tr = np.zeros(nsteps, np.float32)
        rec = None if receivers is None else np.zeros((len(receivers), nsteps), np.float32)
        for s in range(nsteps):
            a, b = self.level(self.i1), self.level(self.i2)
            lib().orc_sem2d_step(self.h, a, b, self.src_element, self.rhs[self.it + s])
            tr[s] = a[self.src_node]
            if rec is not None:
                rec[:, s] = a[np.asarray(receivers)]
            self.i1, self.i2 = self.i2, self.i1
        self.it += nsteps
        return tr, rec

    def current(self):
        return self.level(self.i2)

    def previous(self):
        return self.level(self.i1)

    def coords(self):
        cx = np.zeros(self.nx, np.float32)
        cz = np.zeros(self.nz, np.float32)
        lib().orc_sem2d_coords(self.h, cx, cz)
        return cx, cz

    def mass_y_sh(self):
        M = np.zeros(self.nN, np.float32)
        y = np.zeros(self.nN, np.float32)
        sh = np.zeros(self.nB, np.float32)
        lib().orc_sem2d_get_my(self.h, _vp(M), _vp(y), _vp(sh))
        return M, y, sh

    def boundary_term(self):
        sh = np.zeros(self.nB, np.float32)
        lib().orc_sem2d_boundary_term(self.h, sh)
        return sh

    def apply_k(self, p):
        y = np.zeros(self.nN, np.float32)
        M = np.zeros(self.nN, np.float32)
        lib().orc_sem2d_apply_k(self.h, np.ascontiguousarray(p, np.float32), y, M)
        return y, M


def sem2d_colors(ex, ez):
    """The reference's 4-colour element lists (SEMmesh.cpp:380-429)."""
    mpc = (ex // 2 + ex % 2) * (ez // 2 + ez % 2)
    counts = np.zeros(4, np.int32)
    lists = np.zeros(4 * mpc, np.int32)
    lib().orc_sem2d_colors(ex, ez, counts, lists)
    return [lists[c * mpc:c * mpc + counts[c]] for c in range(4)]


class Sem2d64Problem:
    """The SEM2D step in FP64 on exact box geometry (same mesh, weights,
    source samples and driver loop as Sem2dProblem)."""

    def __init__(self, ex=65, ez=None, lx=2000.0, lz=None, order=3, dt=0.001, f0=10.0, model=None,
                 nsamples=1000, src_element=0, source_order=1):
        ez = ex if ez is None else ez
        lz = lx if lz is None else lz
        nE = ex * ez
        model = np.full(nE, 1500.0) if model is None else np.asarray(model, np.float64)
        dt32 = np.float32(dt)
        self.h = lib().orc_sem2d64_create(ex, ez, lx, lz, order, np.ascontiguousarray(model, np.float64),
                                          float(dt32), int(src_element))
        if not self.h:
            raise ValueError("bad 2D mesh")
        self.nx, self.nz = ex * order + 1, ez * order + 1
        self.nN = self.nx * self.nz
        self.src_node = (src_element % ex) * order + (src_element // ex) * order * self.nx
        self.rhs = np.zeros(nsamples, np.float32)
        lib().orc_source_term(nsamples, dt32, f0, source_order, 0, self.rhs)
        self.p0 = np.zeros(self.nN)
        self.p1 = np.zeros(self.nN)
        self.i1, self.i2 = 0, 1
        self.it = 0

    def __del__(self):
        try:
            lib().orc_sem2d64_destroy(self.h)
        except Exception:
            pass

    def level(self, i):
        return self.p0 if i == 0 else self.p1

    def run(self, nsteps):
        tr = np.zeros(nsteps)
        for s in range(nsteps):
            a, b = self.level(self.i1), self.level(self.i2)
            lib().orc_sem2d64_step(self.h, a, b, float(self.rhs[self.it + s]))
            tr[s] = a[self.src_node]
            self.i1, self.i2 = self.i2, self.i1
        self.it += nsteps
        return tr

    def current(self):
        return self.level(self.i2)
