"""Checks that pin the oracle's physics to statements independent of its code.

Each pin takes the oracle module as an argument, so the same pins run on the
committed oracle (tests/test_oracle.py) and on copies with planted mistakes
(tests/test_oracle_mutations.py), where at least one pin must fail.  None of
them restates the oracle's formulas: the references are

* the Orr-Sommerfeld and Squire linear operators of plane shear flow
  (linearisation of getnonlinear about a mean flow U(y), W(y));
* the Crank-Nicolson form of (D^2 - k^2) v_t = h_v + (D^2 - k^2)^2 v / re,
  omega_t = h_g + (D^2 - k^2) omega / re and U_t = -L_x(0,0) + D^2 U / re,
  assembled from numpy.polynomial.chebyshev derivatives (the solve);
* the Dirichlet eigenvalues of D^2 - k^2 and of the clamped (D^2 - k^2)^2
  pencil (the Crank-Nicolson amplification of one step);
* the Reynolds-stress identity <(omega x u)_x> = d<uv>/dy of an incompressible
  field (the mean-mode Lamb vector);
* the curl of the velocity after a solve (every solved mode divergence-free
  with omega = curl u; omega_x(0,0) = DW, omega_z(0,0) = -DU for the mean).

All grids are small enough that each pin runs in well under a second.
"""
import numpy as np
import numpy.polynomial.chebyshev as C


def pin_params(O, nnx=8, nny=8, nnz=12, dt=1e-2, re=10.0, re_p=0.05):
    """A box where the viscous terms c k^2 (c = dt/2re) are O(1e-2), so a
    wrong operator coefficient is far above rounding."""
    return O.Params(nnx=nnx, nny=nny, nnz=nnz, aphi=1.0, beta=0.5, dt=dt, re=re, re_p=re_p)


def wavenumbers(p):
    """alpha_i = i/aphi and beta_j = j/beta in FFT order (PF:123-134), restated."""
    alpha = np.arange(p.Mx) / p.aphi
    j = np.arange(p.Mz)
    return alpha, np.where(j <= p.Mz // 2, j, j - p.Mz) / p.beta


def points(n):
    return np.cos(np.pi * np.arange(n + 1) / n)


def deriv_matrices(n, rows):
    """Values of T_l and its 1st, 2nd and 4th derivatives at the points y_rows."""
    y = points(n)[rows]
    out = []
    for m in (0, 1, 2, 4):
        mat = np.zeros((len(y), n + 1))
        for l in range(n + 1):
            e = np.zeros(n + 1)
            e[l] = 1.0
            mat[:, l] = C.chebval(y, C.chebder(e, m)) if m else C.chebval(y, e)
        out.append(mat)
    return out


def random_perturbation(p, rng, max_degree):
    """Random v, omega_y coefficients of degree <= max_degree with the layout's
    constraints: zero Nyquist column and (0,0) mode, Hermitian kx = 0 plane."""
    out = []
    for _ in range(2):
        f = rng.standard_normal((p.Ny, p.Mz, p.Mx)) + 1j * rng.standard_normal((p.Ny, p.Mz, p.Mx))
        f[max_degree + 1:] = 0.0
        f[:, p.Mz // 2, :] = 0.0
        f[:, 0, 0] = 0.0
        for j in range(1, p.Mz // 2):
            f[:, p.Mz - j, 0] = np.conj(f[:, j, 0])
        out.append(f)
    return out


def cheb_at_points(coef, n):
    """Chebyshev series along axis 0 evaluated at y_j = cos(pi j/N), axis 0 kept."""
    vals = C.chebval(points(n), coef)                    # (..., n+1)
    return np.moveaxis(vals, -1, 0)


def pin_linearisation(O):
    """getnonlinear is quadratic in the state, so the central difference
    (h(+e) - h(-e)) / 2e about a mean flow (U, W) is exactly its linear part.
    For an incompressible perturbation that part is the Orr-Sommerfeld /
    Squire operator of the mean flow,
        h_v = -i(aU + bW)(D^2 - k^2) v + i(aU'' + bW'') v,
        h_g = -i(aU + bW) omega_y + i(aW' - bU') v,
    which holds here without truncation because U, W have degree 3 and the
    perturbation degree N - 4 (every product stays within degree N)."""
    p = pin_params(O)
    n = p.N
    rng = np.random.default_rng(2024)
    v, oy = random_perturbation(p, rng, n - 4)
    U = np.zeros(p.Ny, complex)
    W = np.zeros(p.Ny, complex)
    U[[0, 1, 2, 3]] = [0.5, 0.2, -0.5, 0.15]
    W[[0, 1, 2, 3]] = [0.1, -0.3, 0.05, 0.2]
    hs = []
    for eps in (1.0, -1.0):
        st = O.zero_state(p)
        st["vnp1"] = eps * v
        st["omegaynp1"] = eps * oy
        st["au0"] = U.copy()
        st["w0"] = W.copy()
        O.kinematics_all(st, p)
        O.getnonlinear(st, p)
        hs.append((st["rhsvn"], st["rhsomegayn"]))
    hv = (hs[0][0] - hs[1][0]) / 2.0
    hg = (hs[0][1] - hs[1][1]) / 2.0
    alpha, beta = wavenumbers(p)
    a3 = alpha[None, None, :]
    b3 = beta[None, :, None]
    k2 = a3 ** 2 + b3 ** 2
    Uy, Wy = C.chebval(points(n), U.real)[:, None, None], C.chebval(points(n), W.real)[:, None, None]
    dU = C.chebval(points(n), C.chebder(U.real))[:, None, None]
    dW = C.chebval(points(n), C.chebder(W.real))[:, None, None]
    d2U = C.chebval(points(n), C.chebder(U.real, 2))[:, None, None]
    d2W = C.chebval(points(n), C.chebder(W.real, 2))[:, None, None]
    vp = cheb_at_points(v, n)
    d2v = cheb_at_points(C.chebder(v, 2), n)
    op = cheb_at_points(oy, n)
    adv = 1j * (a3 * Uy + b3 * Wy)
    hv_ref = -adv * (d2v - k2 * vp) + 1j * (a3 * d2U + b3 * d2W) * vp
    hg_ref = -adv * op + 1j * (a3 * dW - b3 * dU) * vp
    for got, ref, name in ((hv, hv_ref, "h_v"), (hg, hg_ref, "h_g")):
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        assert err < 1e-11, f"linearised {name} differs from the Orr-Sommerfeld/Squire form: {err:.3e}"


def _state_for_solve(O, p, rng):
    st = O.zero_state(p)
    v, oy = random_perturbation(p, rng, p.N - 4)
    st["vnp1"], st["omegaynp1"] = v, oy
    for k in ("rhsvn", "rhsomegayn", "unp1", "wnp1"):       # h^n, h^{n-1} point values
        f = rng.standard_normal((p.Ny, p.Mz, p.Mx)) + 1j * rng.standard_normal((p.Ny, p.Mz, p.Mx))
        st[k] = f
    for k in ("clamb1n", "clamb1nm1", "clamb3n", "clamb3nm1"):
        st[k] = rng.standard_normal(p.Ny) + 0j
    st["au0"] = np.zeros(p.Ny, complex)
    st["au0"][[0, 2, 3]] = [0.5, -0.5, 0.1]
    st["w0"] = np.zeros(p.Ny, complex)
    st["w0"][[1, 2]] = [0.2, -0.05]
    return st


def pin_solve_is_crank_nicolson(O):
    """One solve is the Crank-Nicolson/AB2 step of the PDEs, with operators
    assembled here from numpy Chebyshev derivatives (c = dt/2re):
      v:  [L - cL^2] v' = [L + cL^2] v + dt(3/2 h_v^n - 1/2 h_v^{n-1})  at y_2..y_{N-2},
          v' = Dv' = 0 at y = +-1, with L = D^2 - k^2;
      omega_y: [1 - cL] o' = [1 + cL] o + dt(3/2 h_g^n - 1/2 h_g^{n-1})  at y_1..y_{N-1},
          o' = 0 at y = +-1;
      U (and W with L_z): [1 - cD^2] U' = [1 + cD^2] U - dt(3/2 L_x^n - 1/2 L_x^{n-1})
          at y_1..y_{N-1}, U' = 0 at y = +-1.
    Residuals are measured against the size of the terms."""
    p = pin_params(O)
    n, dt = p.N, p.dt
    c = dt / (2.0 * p.re)
    rng = np.random.default_rng(77)
    st = _state_for_solve(O, p, rng)
    old = O.copy_state(st)
    O.Stepper(p).solve(st)
    alpha, beta = wavenumbers(p)
    iv = np.arange(2, n - 1)
    io = np.arange(1, n)
    M0v, M1v, M2v, M4v = deriv_matrices(n, iv)
    M0o, _, M2o, _ = deriv_matrices(n, io)
    _, M1w, _, _ = deriv_matrices(n, np.array([0, n]))
    B0 = deriv_matrices(n, np.array([0, n]))[0]
    worst = {}
    for j in range(p.Mz):
        if j == p.Mz // 2:
            continue
        for i in range(p.Mx):
            k2 = alpha[i] ** 2 + beta[j] ** 2
            if k2 == 0.0:
                continue
            L = M2v - k2 * M0v
            L2 = M4v - 2 * k2 * M2v + k2 ** 2 * M0v
            vn, vo = st["vnp1"][:, j, i], old["vnp1"][:, j, i]
            f = dt * (1.5 * old["rhsvn"][iv, j, i] - 0.5 * old["unp1"][iv, j, i])
            lhs = (L - c * L2) @ vn
            rhs = (L + c * L2) @ vo + f
            scale = np.max(np.abs(rhs)) + np.max(np.abs(lhs))
            worst[("v", i, j)] = np.max(np.abs(lhs - rhs)) / scale
            bc = np.max(np.abs(np.concatenate([B0 @ vn, M1w @ vn]))) / np.max(np.abs(vn))
            worst[("v-bc", i, j)] = bc
            Lo = M2o - k2 * M0o
            on, oo = st["omegaynp1"][:, j, i], old["omegaynp1"][:, j, i]
            f = dt * (1.5 * old["rhsomegayn"][io, j, i] - 0.5 * old["wnp1"][io, j, i])
            lhs = (M0o - c * Lo) @ on
            rhs = (M0o + c * Lo) @ oo + f
            scale = np.max(np.abs(rhs)) + np.max(np.abs(lhs))
            worst[("omega_y", i, j)] = np.max(np.abs(lhs - rhs)) / scale
            worst[("omega_y-bc", i, j)] = np.max(np.abs(B0 @ on)) / np.max(np.abs(on))
    for mean, ln, lnm1 in (("au0", "clamb1n", "clamb1nm1"), ("w0", "clamb3n", "clamb3nm1")):
        xn, xo = st[mean], old[mean]
        f = -dt * (1.5 * old[ln][io] - 0.5 * old[lnm1][io])
        lhs = (M0o - c * M2o) @ xn
        rhs = (M0o + c * M2o) @ xo + f
        scale = np.max(np.abs(rhs)) + np.max(np.abs(lhs))
        worst[(mean,)] = np.max(np.abs(lhs - rhs)) / scale
        worst[(mean + "-bc",)] = np.max(np.abs(B0 @ xn)) / np.max(np.abs(xn))
    key = max(worst, key=worst.get)
    assert worst[key] < 1e-10, f"solve is not the Crank-Nicolson step at {key}: {worst[key]:.3e}"


def pin_crank_nicolson_eigenmodes(O):
    """An eigenfunction of the Dirichlet problem (D^2 - k^2) phi = lam phi
    (omega_y), or of the clamped pencil (D^2 - k^2)^2 phi = lam (D^2 - k^2) phi
    with phi = phi' = 0 at the walls (v), is multiplied in one unforced solve
    by (1 + c lam)/(1 - c lam).  lam comes from an eigenproblem assembled here
    from numpy derivatives, not from the oracle's operators."""
    import scipy.linalg
    p = pin_params(O, nnz=20, dt=2e-3, re=50.0, re_p=0.0)
    n = p.N
    c = p.dt / (2.0 * p.re)
    alpha, beta = wavenumbers(p)
    i, j = 1, 1
    k2 = alpha[i] ** 2 + beta[j] ** 2
    ends = np.array([0, n])
    E0, E1, _, _ = deriv_matrices(n, ends)
    for which in ("omegay", "v"):
        if which == "omegay":
            rows = np.arange(1, n)
            M0, _, M2, _ = deriv_matrices(n, rows)
            A = np.vstack([M2 - k2 * M0, E0])
            B = np.vstack([M0, np.zeros_like(E0)])
        else:
            rows = np.arange(2, n - 1)
            M0, _, M2, M4 = deriv_matrices(n, rows)
            A = np.vstack([M4 - 2 * k2 * M2 + k2 ** 2 * M0, E0, E1])
            B = np.vstack([M2 - k2 * M0, np.zeros_like(E0), np.zeros_like(E1)])
        lam, vec = scipy.linalg.eig(A, B)
        ok = np.isfinite(lam) & (np.abs(lam) < 1e6)
        lam, vec = lam[ok], vec[:, ok]
        order = np.argsort(-lam.real)                       # least damped first
        for k in order[:3]:
            mode = vec[:, k] / np.max(np.abs(vec[:, k])) * 1e-9
            st = O.zero_state(p)
            key = "omegaynp1" if which == "omegay" else "vnp1"
            st[key][:, j, i] = mode
            O.Stepper(p).solve(st)
            g = (1.0 + c * lam[k]) / (1.0 - c * lam[k])
            err = np.max(np.abs(st[key][:, j, i] - g * mode)) / np.max(np.abs(mode))
            assert err < 1e-9, f"{which}: eigenmode lam={lam[k]:.6g} not scaled by the CN factor: {err:.3e}"


def full_spectrum(f, p):
    """(kx, kz) -> Chebyshev coefficients of the real field the padded
    transforms see (Nyquist dropped, kx = 0 plane Hermitian-projected)."""
    h = p.Mz // 2
    out = {}
    for j in range(p.Mz):
        if j == h:
            continue
        kz = j if j < h else j - p.Mz
        for i in range(p.Mx):
            if i == 0:
                jn = (-j) % p.Mz
                col = 0.5 * (f[:, j, 0] + np.conj(f[:, jn, 0]))
            else:
                col = f[:, j, i]
            out[(i, kz)] = col
            if i > 0:
                out[(-i, -kz)] = np.conj(col)
    return out


def plane_mean_product(a, b, p):
    """Plane average <ab>(y) of two real fields as an exact Chebyshev series
    (degree 2N): sum over all modes of a_k conj(b_k) (Parseval)."""
    sa, sb = full_spectrum(a, p), full_spectrum(b, p)
    acc = np.zeros(2 * p.N + 1)
    for key, ca in sa.items():
        cb = sb[key]
        prod = C.chebmul(ca.real, cb.real) + C.chebmul(ca.imag, cb.imag)
        acc[:len(prod)] += prod
    return acc


def pin_mean_lamb(O):
    """Mean-mode Lamb vector of an incompressible field: with
    omega x u = (u.grad)u - grad|u|^2/2 the plane average of the x and z
    components is d<uv>/dy and d<vw>/dy (the x, z derivatives average out), so
    L_x(0,0) = d<uv>/dy - re_p and L_z(0,0) = d<vw>/dy, truncated to degree N."""
    p = pin_params(O, nnz=10)
    rng = np.random.default_rng(5)
    v, oy = random_perturbation(p, rng, p.N)
    st = O.zero_state(p)
    st["vnp1"], st["omegaynp1"] = v, oy
    st["au0"][[0, 1, 2]] = [0.5, 0.1, -0.5]
    st["w0"][[0, 2]] = [0.2, -0.2]
    O.kinematics_all(st, p)
    u, w, vv = st["unp1"].copy(), st["wnp1"].copy(), st["vnp1"].copy()
    O.getnonlinear(st, p)
    n = p.N
    for name, a, b, shift in (("clamb1n", u, vv, p.re_p), ("clamb3n", vv, w, 0.0)):
        ref = C.chebder(plane_mean_product(a, b, p))[:n + 1].copy()
        ref[0] -= shift
        got = st[name]                                    # point values y_0..y_N
        ref_pts = C.chebval(points(n), ref)
        err = np.max(np.abs(got - ref_pts)) / np.max(np.abs(ref_pts))
        assert err < 1e-11, f"{name}: mean Lamb vector is not the Reynolds-stress gradient: {err:.3e}"


def pin_mean_curl(O):
    """After a solve, the mean-mode vorticity is the curl of the mean flow:
    omega_x(0,0) = DW, omega_z(0,0) = -DU (omega = curl (U(y), 0, W(y)))."""
    p = pin_params(O)
    rng = np.random.default_rng(31)
    st = _state_for_solve(O, p, rng)
    O.Stepper(p).solve(st)
    for vort, mean, sign in (("omegaxnp1", "w0", 1.0), ("omegaznp1", "au0", -1.0)):
        ref = np.zeros(p.Ny, complex)
        d = C.chebder(st[mean])
        ref[:len(d)] = sign * d
        got = st[vort][:, 0, 0]
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, f"{vort}(0,0) is not the curl of the mean flow: {err:.3e}"
    for vel, mean in (("unp1", "au0"), ("wnp1", "w0")):
        assert np.array_equal(st[vel][:, 0, 0], st[mean]), vel


def pin_solved_kinematics(O):
    """After a solve every solved mode (k^2 != 0, not the span Nyquist) is an
    incompressible velocity with its curl: i a u + Dv + i b w = 0,
    omega_y = i b u - i a w, omega_x = Dw - i b v, omega_z = i a v - Du, with D
    numpy's Chebyshev derivative."""
    p = pin_params(O)
    rng = np.random.default_rng(8)
    st = _state_for_solve(O, p, rng)
    O.Stepper(p).solve(st)
    alpha, beta = wavenumbers(p)

    def d(f):
        out = np.zeros_like(f)
        g = C.chebder(f)
        out[:len(g)] = g
        return out

    worst = 0.0
    for j in range(p.Mz):
        if j == p.Mz // 2:
            continue
        for i in range(p.Mx):
            a, b = alpha[i], beta[j]
            if a * a + b * b == 0.0:
                continue
            u, v, w = st["unp1"][:, j, i], st["vnp1"][:, j, i], st["wnp1"][:, j, i]
            ox, oy, oz = st["omegaxnp1"][:, j, i], st["omegaynp1"][:, j, i], st["omegaznp1"][:, j, i]
            scale = max(np.max(np.abs(d(v))), np.max(np.abs(a * u)), np.max(np.abs(b * w)), np.max(np.abs(oy)))
            for res in (1j * a * u + d(v) + 1j * b * w, 1j * b * u - 1j * a * w - oy,
                        d(w) - 1j * b * v - ox, 1j * a * v - d(u) - oz):
                worst = max(worst, np.max(np.abs(res)) / scale)
    assert worst < 1e-12, f"solved modes are not a divergence-free velocity with its curl: {worst:.3e}"


ALL_PINS = (pin_linearisation, pin_solved_kinematics, pin_solve_is_crank_nicolson, pin_crank_nicolson_eigenmodes, pin_mean_lamb,
            pin_mean_curl)
