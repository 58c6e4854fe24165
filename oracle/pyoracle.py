"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
cpu_baseline leg as the checker.  The product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_mutant(k, out_dir):
    """Builds the oracle with mistake k planted (ORACLE_MUTANT=k, see the
    header of simple_oracle.hpp) into out_dir; returns the library path."""
    path = os.path.join(out_dir, f"liboracle_m{k}.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-shared", f"-DORACLE_MUTANT={k}", "-o", path,
                    os.path.join(_HERE, "oracle_capi.cpp")], check=True)
    return path


def lib(types, path=None):
    """Load liboracle.so (or the library at `path`, e.g. a mutant); `types`
    is the package's _types module (shared struct mirrors).  Builds the
    default library first when only the sources are present."""
    if path is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
    if path in _LIBS:
        return _LIBS[path]
    L = C.CDLL(path)
    P = C.POINTER
    dp = P(C.c_double)
    ip = P(C.c_int)
    L.orc_simple_create.restype = C.c_void_p
    L.orc_simple_create.argtypes = [C.c_int, P(types.SimpleConfig), dp, dp]
    L.orc_simple_destroy.argtypes = [C.c_void_p]
    L.orc_simple_step.argtypes = [C.c_void_p, C.c_int, P(types.StepStats), dp]
    L.orc_simple_run.argtypes = [C.c_void_p, P(types.StepStats), dp, dp, C.c_int]
    L.orc_simple_plane.argtypes = [C.c_void_p, C.c_double, P(types.Patch)]
    L.orc_simple_get_field.argtypes = [C.c_void_p, C.c_int, dp]
    L.orc_simple_set_field.argtypes = [C.c_void_p, C.c_int, dp]
    L.orc_simple_boundary_patch.argtypes = [C.c_void_p, C.c_int, P(types.Patch)]
    L.orc_simple_apply_bc.argtypes = [C.c_void_p]
    L.orc_simple_assemble.argtypes = [C.c_void_p, C.c_int]
    L.orc_simple_csr_size.argtypes = [C.c_void_p, C.c_int, ip, ip]
    L.orc_simple_csr_get.argtypes = [C.c_void_p, C.c_int, ip, ip, dp, dp, dp, ip]
    L.orc_simple_residuals.argtypes = [C.c_void_p, dp]
    L.orc_simple_face.restype = C.c_double
    L.orc_simple_face.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.orc_simple_sou.restype = C.c_double
    L.orc_simple_sou.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_double] * 4
    for f in (L.orc_jacobi_csr, L.orc_pcg_csr):
        f.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int]
    L.orc_thermal_create.restype = C.c_void_p
    L.orc_thermal_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp, ip, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_double, dp, dp, dp, dp]
    L.orc_thermal_destroy.argtypes = [C.c_void_p]
    L.orc_thermal_csr_size.argtypes = [C.c_void_p, ip, ip]
    L.orc_thermal_csr_get.argtypes = [C.c_void_p, ip, ip, dp, dp]
    L.orc_thermal_solve.argtypes = [C.c_void_p, dp, C.c_double, C.c_int, dp, dp]
    _LIBS[path] = L
    return L


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class OracleSimple:
    """The reference SIMPLE solver restated on the CPU (prec 64 or 32)."""

    def __init__(self, types, cfg, gamma, inlet=None, prec=64, lib_path=None):
        self.T = types
        self.L = lib(types, lib_path)
        self.cfg = cfg.copy()
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        assert g.shape == (cfg.ny, cfg.nx)
        inl = None if inlet is None else np.ascontiguousarray(inlet, dtype=np.float64)
        if inl is not None:
            assert inl.shape == (cfg.ny + 2,)
        self.h = self.L.orc_simple_create(prec, C.byref(self.cfg), _d(g), None if inl is None else _d(inl))
        self.prec = prec

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_simple_destroy(self.h)
            self.h = None

    def shape(self, name):
        return self.T.field_shape(name, self.cfg.nx, self.cfg.ny)

    def get(self, name):
        out = np.zeros(self.shape(name))
        self.L.orc_simple_get_field(self.h, self.T.FIELDS[name], _d(out))
        return out

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape(name)
        self.L.orc_simple_set_field(self.h, self.T.FIELDS[name], _d(a))

    def step(self, n=1):
        st = self.T.StepStats()
        secs = C.c_double()
        self.L.orc_simple_step(self.h, n, C.byref(st), C.byref(secs))
        self.last_seconds = secs.value
        return st.as_dict()

    def run(self, history_len=0):
        st = self.T.StepStats()
        hist = np.zeros(max(history_len, 1))
        dph = np.zeros(max(history_len, 1))
        it = self.L.orc_simple_run(self.h, C.byref(st), _d(hist), _d(dph), history_len)
        d = st.as_dict()
        d["iterations"] = it
        d["history"] = hist[: min(it, history_len)]
        d["dp_history"] = dph[: min(it, history_len)]
        return d

    def plane(self, x):
        m = self.T.Patch()
        self.L.orc_simple_plane(self.h, float(x), C.byref(m))
        return m.as_dict()

    def patch(self, inlet):
        m = self.T.Patch()
        self.L.orc_simple_boundary_patch(self.h, 1 if inlet else 0, C.byref(m))
        return m.as_dict()

    def dp_static(self):
        return self.patch(True)["avg_static"] - self.patch(False)["avg_static"]

    def apply_bc(self):
        self.L.orc_simple_apply_bc(self.h)

    def assemble(self, which):
        """which: 'u', 'v' or 'p'; returns the CSR system as a dict."""
        w = {"u": 0, "v": 1, "p": 2}[which]
        self.L.orc_simple_assemble(self.h, w)
        n, nnz = C.c_int(), C.c_int()
        self.L.orc_simple_csr_size(self.h, w, C.byref(n), C.byref(nnz))
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        rhs = np.zeros(n.value)
        sol = np.zeros(n.value)
        ij = np.zeros(2 * n.value, np.int32)
        self.L.orc_simple_csr_get(self.h, w, _i(rp), _i(ci), _d(val), _d(rhs), _d(sol), _i(ij))
        return dict(row_ptr=rp, col_idx=ci, values=val, rhs=rhs, sol=sol, ij=ij.reshape(-1, 2))

    def residual_fields(self, u, v, p):
        """Residuals of the discrete equations a converged SIMPLE state
        satisfies, at the given fields: U / V momentum A x - b at x = the
        field (relaxation cancels) and the mass residual of the pressure
        DOFs, scattered into arrays of the u, v, p shapes (zero elsewhere).
        The fields of this object are overwritten."""
        import scipy.sparse as sp
        self.set("u", u)
        self.set("v", v)
        self.set("p", p)
        self.apply_bc()
        self.set("uStar", self.get("u"))
        self.set("vStar", self.get("v"))
        out = {}
        for w in ("u", "v"):
            sysd = self.assemble(w)
            m = len(sysd["rhs"])
            A = sp.csr_matrix((sysd["values"], sysd["col_idx"], sysd["row_ptr"]), shape=(m, m))
            r = np.zeros(self.shape(w))
            r[sysd["ij"][:, 0], sysd["ij"][:, 1]] = A @ sysd["sol"] - sysd["rhs"]
            out[w] = r
        sysd = self.assemble("p")
        r = np.zeros(self.shape("p"))
        r[sysd["ij"][:, 0], sysd["ij"][:, 1]] = sysd["rhs"]
        out["p"] = r
        return out

    def residuals(self):
        out = np.zeros(6)
        self.L.orc_simple_residuals(self.h, _d(out))
        return out

    def face(self, kind, i, j):
        k = ["alphaAtU", "alphaAtV", "gammaAtU", "gammaAtV", "southU", "northU", "westV", "eastV"].index(kind)
        return self.L.orc_simple_face(self.h, k, i, j)

    def sou(self, is_u, i, j, Fe, Fw, Fn, Fs):
        return self.L.orc_simple_sou(self.h, 1 if is_u else 0, i, j, Fe, Fw, Fn, Fs)


def jacobi_csr(types, n, rp, ci, val, b, x0, tol, maxit, prec=64):
    L = lib(types)
    x = np.array(x0, dtype=np.float64)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    it = L.orc_jacobi_csr(prec, n, _i(rp), _i(ci), _d(val), _d(b), _d(x), tol, maxit)
    return x, it


def pcg_csr(types, n, rp, ci, val, b, x0, tol, maxit, prec=64):
    L = lib(types)
    x = np.array(x0, dtype=np.float64)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    it = L.orc_pcg_csr(prec, n, _i(rp), _i(ci), _d(val), _d(b), _d(x), tol, maxit)
    return x, it


class OracleThermal:
    """Reference thermal assembly (FP64) + CPU BiCGStab(Jacobi)."""

    def __init__(self, types, stack, gamma, u, v):
        self.L = lib(types)
        nx, ny = stack["nx"], stack["ny"]
        self.nx, self.ny, self.nz = nx, ny, len(stack["dz"])
        self.keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                     (stack["dz"], gamma, u, v, stack["q0"])]
        ch = np.ascontiguousarray(stack["channel"], dtype=np.int32)
        self.keep.append(ch)
        dz, g, uu, vv, q0 = self.keep[:5]
        for a in (g, uu, vv, q0):
            assert a.shape == (ny, nx)
        self.h = self.L.orc_thermal_create(nx, ny, self.nz, stack.get("dx", 2e-5), stack.get("dy", 2e-5), _d(dz),
                                           _i(ch), stack["ht"], stack["k_s"], stack["k_f"], stack["rho"],
                                           stack["cp"], stack["t_in"], stack.get("k_mode", 1), stack["qk"],
                                           _d(g), _d(uu), _d(vv), _d(q0))
        self.t_in = stack["t_in"]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_thermal_destroy(self.h)
            self.h = None

    def csr(self):
        n, nnz = C.c_int(), C.c_int()
        self.L.orc_thermal_csr_size(self.h, C.byref(n), C.byref(nnz))
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        b = np.zeros(n.value)
        self.L.orc_thermal_csr_get(self.h, _i(rp), _i(ci), _d(val), _d(b))
        return rp, ci, val, b

    def solve(self, rtol=1e-10, maxit=20000, T0=None):
        T = np.full(self.nx * self.ny * self.nz, self.t_in) if T0 is None else np.array(T0, dtype=np.float64)
        rr, secs = C.c_double(), C.c_double()
        it = self.L.orc_thermal_solve(self.h, _d(T), rtol, maxit, C.byref(rr), C.byref(secs))
        self.last_seconds = secs.value
        self.last_relres = rr.value
        self.last_iters = it
        return T.reshape(self.nz, self.ny, self.nx)

    def matrix(self):
        """Assembled system as a scipy CSR matrix and b."""
        import scipy.sparse as sp
        rp, ci, val, b = self.csr()
        n = len(b)
        return sp.csr_matrix((val, ci, rp), shape=(n, n)), b

    def direct(self, transpose=False, rhs=None):
        """Sparse LU solve of A T = b (or A^T x = rhs): the exact solution of
        the oracle-assembled system, shape (nz, ny, nx)."""
        import scipy.sparse.linalg as sla
        A, b = self.matrix()
        if transpose:
            A = A.T.tocsc()
        r = b if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64).ravel()
        return sla.spsolve(A.tocsc(), r).reshape(self.nz, self.ny, self.nx)

    def rel_residual(self, T):
        A, b = self.matrix()
        return np.linalg.norm(b - A @ np.ravel(T)) / np.linalg.norm(b)


def grad_magnitude(T, dx, dy, dz):
    """|grad T| per cell (compute_gradient_components, thermal_output.cpp:
    63-120): (T[i+1] - T[i-1]) / (2 h) inside, first differences at the
    faces; in z the central difference divides by 2 * (dz[k-1] + dz[k]) / 2
    and the face differences by dz[0] / dz[nz-1].  An axis of one cell
    contributes 0."""
    nz, ny, nx = T.shape
    g2 = np.zeros_like(T)
    for ax, h in ((2, dx), (1, dy)):
        if T.shape[ax] > 1:
            g2 += np.gradient(T, h, axis=ax) ** 2
    if nz > 1:
        gz = np.empty_like(T)
        gz[0] = (T[1] - T[0]) / dz[0]
        gz[-1] = (T[-1] - T[-2]) / dz[-1]
        for k in range(1, nz - 1):
            gz[k] = (T[k + 1] - T[k - 1]) / (2.0 * (0.5 * (dz[k - 1] + dz[k])))
        g2 += gz ** 2
    return np.sqrt(g2)


def thermal_metrics(stack, gamma, u, T, pnorm_exp=10.0):
    """numpy restatement of compute_thermal_metrics (thermal_output.cpp:135-
    321, including the gradient pass), Tb_max / Tb_pnorm of
    compute_optimization_results (thermal_solver.cpp:361-381) and the energy
    balance of ThermalSolver3D::solve (thermal_solver.cpp:592-620).  Base
    layers are those below the first channel layer (k < nz_solid)."""
    dz = np.asarray(stack["dz"], dtype=np.float64)
    ch = np.asarray(stack["channel"]) != 0
    nz, ny, nx = T.shape
    dx, dy = stack["dx"], stack["dy"]
    k0 = int(np.argmax(ch)) if ch.any() else nz
    V = (dx * dy * dz)[:, None, None] * np.ones((nz, ny, nx))
    base = np.zeros((nz, ny, nx), bool)
    base[:k0] = True
    fl = ch[:, None, None] & (np.asarray(gamma)[None] > 0.01)
    solid = ~base & ~fl

    def avg(mask):
        return (T[mask] * V[mask]).sum() / V[mask].sum() if mask.any() else 0.0

    def std(mask, a):
        return np.sqrt((V[mask] * (T[mask] - a) ** 2).sum() / V[mask].sum()) if mask.any() else 0.0

    def rng(mask):
        return T[mask].max() - T[mask].min() if mask.any() else 0.0

    allm = np.ones_like(base)
    m = dict(t_avg_base=avg(base), t_avg_fluid=avg(fl), t_avg_solid=avg(solid), t_avg_global=avg(allm))
    m.update(t_std_base=std(base, m["t_avg_base"]), t_std_fluid=std(fl, m["t_avg_fluid"]),
             t_std_global=std(allm, m["t_avg_global"]))
    m.update(t_range_base=rng(base), t_range_fluid=rng(fl), t_range_global=rng(allm))

    def unif(sd, a):
        return 0.0 if a <= 0 else max(0.0, min(1.0, 1.0 - sd / a))

    m["t_uniformity_base"] = unif(m["t_std_base"], m["t_avg_base"])
    m["t_uniformity_global"] = unif(m["t_std_global"], m["t_avg_global"])
    m["tb_max"] = T[0].max()
    m["tb_pnorm"] = np.mean(T[0] ** pnorm_exp) ** (1.0 / pnorm_exp)
    m["q_in"] = (np.asarray(stack["q0"]) * dx * dy).sum()
    # outlet: i = nx - 1, fluid channel cells with u_xy * pois > 0
    mdot = mdotT = 0.0
    for k in range(nz):
        if not ch[k]:
            continue
        zb = dz[k0:k].sum()
        z = (zb + 0.5 * dz[k]) / stack["ht"]
        po = 4 * z * (1 - z)
        uo = np.asarray(u)[:, nx - 1] * po
        sel = (np.asarray(gamma)[:, nx - 1] > 0.01) & (uo > 0)
        mm = stack["rho"] * uo[sel] * dy * dz[k]
        mdot += mm.sum()
        mdotT += (mm * T[k, sel, nx - 1]).sum()
    m["t_out"] = mdotT / mdot if mdot > 0 else stack["t_in"]
    m["q_out"] = mdot * stack["cp"] * (m["t_out"] - stack["t_in"])
    m["imbalance_pct"] = 100 * (m["q_in"] - m["q_out"]) / (m["q_in"] + 1e-30)
    # gradient pass (thermal_output.cpp:270-321)
    G = grad_magnitude(T, dx, dy, dz)
    q = int(np.argmax(G))  # first maximum in (k, j, i) order
    m["grad_max"] = G.ravel()[q]
    m["grad_max_k"], m["grad_max_j"], m["grad_max_i"] = np.unravel_index(q, G.shape)
    m["grad_avg"] = (G * V).sum() / V.sum()
    m["grad_max_base"] = G[base].max() if base.any() else -1e30
    m["grad_avg_base"] = (G[base] * V[base]).sum() / V[base].sum() if base.any() else 0.0
    return m


def thermal_objective(stack, gamma, u, T, objective, pnorm_exp=8.0):
    m = thermal_metrics(stack, gamma, u, T, pnorm_exp)
    return m["t_avg_base"] if objective == "t_avg_base" else m["tb_pnorm"]


def surrogate_adjoint(k, P, T=None):
    """2-D conduction surrogate adjoint of AdjointMethod/thermal_adjoint.py:
    8-159 restated with scipy sparse: K with harmonic-mean faces k_eff / dx^2
    (dx = 1 / nx) and 2 k / dx^2 per boundary side on the diagonal (:45-76);
    T = K^-1 P unless given (:81-88); lambda = K^-T 1/n (:93-96); sens =
    -sum_nb (l_i - l_nb) 2 k_nb^2 / (k_i + k_nb)^2 / dx^2 (T_i - T_nb)
    - l_i 2 (boundary sides) / dx^2 T_i (:101-123).  Returns T, lambda, sens."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    k = np.asarray(k, dtype=np.float64)
    ny, nx = k.shape
    n = nx * ny
    dx = 1.0 / nx
    idx = np.arange(n).reshape(ny, nx)
    rows, cols, vals = [], [], []
    diag = np.zeros((ny, nx))
    for di, dj in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        i0, i1 = max(0, -di), ny - max(0, di)
        j0, j1 = max(0, -dj), nx - max(0, dj)
        ki = k[i0:i1, j0:j1]
        kn = k[i0 + di:i1 + di, j0 + dj:j1 + dj]
        c = 2 * ki * kn / (ki + kn) / dx ** 2
        rows.append(idx[i0:i1, j0:j1].ravel())
        cols.append(idx[i0 + di:i1 + di, j0 + dj:j1 + dj].ravel())
        vals.append(-c.ravel())
        diag[i0:i1, j0:j1] += c
    bc = np.zeros((ny, nx))
    bc[:, 0] += 1
    bc[:, -1] += 1
    bc[0, :] += 1
    bc[-1, :] += 1
    diag += 2 * bc * k / dx ** 2
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    if T is None:
        T = sla.spsolve(K.tocsc(), np.asarray(P, dtype=np.float64).ravel()).reshape(ny, nx)
    T = np.asarray(T, dtype=np.float64)
    lam = sla.spsolve(K.T.tocsc(), np.ones(n) / n).reshape(ny, nx)
    sens = np.zeros((ny, nx))
    for di, dj in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        i0, i1 = max(0, -di), ny - max(0, di)
        j0, j1 = max(0, -dj), nx - max(0, dj)
        ki = k[i0:i1, j0:j1]
        kn = k[i0 + di:i1 + di, j0 + dj:j1 + dj]
        dkeff = 2 * kn ** 2 / (ki + kn) ** 2
        sens[i0:i1, j0:j1] += -(lam[i0:i1, j0:j1] - lam[i0 + di:i1 + di, j0 + dj:j1 + dj]) * (dkeff / dx ** 2) * \
            (T[i0:i1, j0:j1] - T[i0 + di:i1 + di, j0 + dj:j1 + dj])
    sens += -lam * (2 * bc / dx ** 2) * T
    return T, lam, sens
