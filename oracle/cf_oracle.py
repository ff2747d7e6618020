"""CPU oracle for the per-timestep core of the Fourier-Fourier-Chebyshev channel DNS.

TEST INFRASTRUCTURE ONLY.  This module is a numpy restatement of the reference
Fortran solver (``poiseuilleflow.f90`` = PF, ``paraandcons.f90`` = PC) at one
MPI rank (P=1).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline.  The product path (``channel_flow_amd``) never calls
it.

Every function cites the reference lines it restates.  The reference itself
cannot be compiled here (it links FFTW 2.1.5, which is not in this image), so
the restatement is pinned by closed forms and invariants in
``tests/test_oracle.py``:

* ``scheb``                  -> numpy.polynomial.chebyshev derivatives at the points
* ``dz``                     -> numpy.polynomial.chebyshev.chebder
* ``fourier_to_phys``        -> brute-force Fourier series at the physical points
* ``phys_to_fourier``        -> inverse of fourier_to_phys on the retained band
* ``cheb_padded_product``    -> numpy.polynomial.chebyshev.chebmul truncated to degree N
* ``getlamb``                -> brute-force 3-D truncated convolution (tiny grids)
* operators / ``solve``      -> A.A^-1 = I (the reference's own PF:206-212 check),
                                laminar fixed point, Crank-Nicolson decay factors,
                                kinematic identities (divergence, omega_y = curl)
* ``gsij``                   -> numpy.linalg.solve
* whole step                 -> Orr-Sommerfeld Tollmien-Schlichting eigenvalue
                                (Orszag 1971) at infinitesimal amplitude

Array conventions (identical in memory to the reference's real*8 arrays):

* spectral field ``(0:nx,0:ny,0:nz)`` with Re/Im interleaved along dim 1
  (PF:253-263) == C-order ``complex128[Ny][Mz][Mx]``:
  axis 0 = Chebyshev index k (or wall-normal point), axis 1 = spanwise mode in
  FFT order (index Mz/2 is the zeroed Nyquist), axis 2 = streamwise mode 0..Mx-1;
* mean-mode arrays ``(2,0:nz)`` == ``complex128[Ny]``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# --------------------------------------------------------------------------
# parameters and grids (PC:4-26, PF:120-148)
# --------------------------------------------------------------------------


@dataclasses.dataclass
class Params:
    """Subset of PC:4-26 that the per-timestep core reads (P=1)."""

    nnx: int        # streamwise grid points (PC:5); Mx = nnx/2 modes are kept
    nny: int        # spanwise grid points (PC:6); Mz = nny modes incl. Nyquist
    nnz: int        # Chebyshev degree N (PC:7); Ny = N+1 coefficients
    aphi: float     # Lx = 2*pi*aphi (PC:21)
    beta: float     # Lz = 2*pi*beta (PC:21)
    dt: float       # time step (PC:23)
    re: float       # Reynolds number (PC:24)
    re_p: float     # mean pressure gradient (PC:25)

    @property
    def Mx(self):
        return self.nnx // 2

    @property
    def Mz(self):
        return self.nny

    @property
    def N(self):
        return self.nnz

    @property
    def Ny(self):
        return self.nnz + 1

    @property
    def Lx(self):   # padded streamwise length 3*nnx/2 (plan9/10, PF:108-109)
        return 3 * self.nnx // 2

    @property
    def Lz(self):   # padded spanwise length 3*nny/2 (plan8/11, PF:107,110)
        return 3 * self.nny // 2


def wavenumbers(p: Params):
    """alpha_i (PF:132-134, id=0) and beta_j in FFT order (PF:123-130)."""
    alpha = np.arange(p.Mx, dtype=np.float64) / p.aphi
    j = np.arange(p.Mz)
    beta = np.where(j <= p.Mz // 2, j, j - p.Mz).astype(np.float64) / p.beta
    return alpha, beta


def k2_grid(p: Params):
    alpha, beta = wavenumbers(p)
    return alpha[None, :] ** 2 + beta[:, None] ** 2        # (Mz, Mx)


def scheb(n: int):
    """T0, T1, T2, T4: T_l^(m)(y_j) at y_j = cos(pi j/n) (SCHEB, PF:3272-3342).

    Rows j are points, columns l are degrees.  T4 is left 0 at the two walls,
    where the reference never reads it (PF:3357-3368 uses rows 2..n-2).
    """
    t0 = np.zeros((n + 1, n + 1))
    t1 = np.zeros_like(t0)
    t2 = np.zeros_like(t0)
    t4 = np.zeros_like(t0)
    lv = np.arange(n + 1, dtype=np.float64)
    sgn = np.where(np.arange(n + 1) % 2 == 0, 1.0, -1.0)
    t0[0, :] = 1.0                                        # PF:3297
    t0[n, :] = sgn                                        # PF:3298
    t1[0, :] = lv ** 2                                    # PF:3299
    t1[n, :] = -sgn * lv ** 2                             # PF:3300
    t2[0, :] = lv ** 2 * (lv ** 2 - 1.0) / 3.0            # PF:3301
    t2[n, :] = sgn * lv ** 2 * (lv ** 2 - 1.0) / 3.0      # PF:3302
    h0 = 4.0 * math.atan(1.0) / n
    for j in range(1, n):                                 # PF:3305-3338
        q = h0 * j
        c1 = math.cos(q)
        c2 = c1 * c1
        s1 = math.sin(q)
        s2 = s1 * s1
        ds1 = 1.0 / s1
        ds3 = ds1 / s2
        ds7 = ds3 / s2 / s2
        qj = q * lv
        sj = np.sin(qj)
        cj = np.cos(qj)
        bs = lv * lv * s2
        bc = lv * cj * s1
        t0[j] = cj
        t1[j] = lv * sj * ds1
        t2[j] = lv * (sj * c1 - bc) * ds3
        t4[j] = lv * (sj * c1 * (9.0 + 6.0 * (c2 - bs)) - bc * (4.0 + 11.0 * c2 - bs)) * ds7
    return t0, t1, t2, t4


def dz(a: np.ndarray) -> np.ndarray:
    """Chebyshev coefficient derivative along axis 0 (dz, PF:570-599).

    b_i = (2/c_i) sum_{j>i, i+j odd} j a_j with c_0 = 2, evaluated as the
    reference's O(n^2) double loop (vectorised over trailing axes).
    """
    n = a.shape[0] - 1
    b = np.zeros_like(a)
    for i in range(n):
        fac = 1.0 if i == 0 else 2.0
        for j in range(i + 1, n + 1):
            if (i + j) % 2:
                b[i] = b[i] + fac * float(j) * a[j]
    return b


# --------------------------------------------------------------------------
# 3/2-padded Fourier transforms (ftrthreetotwo PF:1024-1185 with expand
# PF:721-765, getall PF:1350-1362; transposes/equipment are identities at P=1)
# --------------------------------------------------------------------------


def fourier_to_phys(f: np.ndarray, p: Params) -> np.ndarray:
    """ftrthreetotwo(+1): complex (K, Mz, Mx) -> real (K, Lz, Lx) on the 3/2 grid.

    Span: zero the Nyquist (PF:730), pad (PF:732-744), unnormalised backward
    FFT (plan8, PF:1073).  Stream: Hermitian completion (getall), pad,
    backward FFT (plan9) and keep the real part (PF:1116-1118), which projects
    the kx=0 plane onto its Hermitian part.
    """
    K = f.shape[0]
    h = p.Mz // 2
    g = np.zeros((K, p.Lz, p.Mx), dtype=np.complex128)
    g[:, :h] = f[:, :h]
    if h > 1:
        g[:, p.Lz - (h - 1):] = f[:, h + 1:]
    hz = np.fft.ifft(g, axis=1) * p.Lz
    x = np.zeros((K, p.Lz, p.Lx), dtype=np.complex128)
    x[..., :p.Mx] = hz
    if p.Mx > 1:
        x[..., p.Lx - (p.Mx - 1):] = np.conj(hz[..., 1:][..., ::-1])
    return (np.fft.ifft(x, axis=2) * p.Lx).real


def phys_to_fourier(r: np.ndarray, p: Params) -> np.ndarray:
    """ftrthreetotwo(-1): real (K, Lz, Lx) -> complex (K, Mz, Mx).

    Forward stream FFT keeping modes 0..Mx-1 (plan10, PF:1135-1142), forward
    span FFT (plan11, PF:1163), truncation that keeps slot Mz/2 (expand(-1),
    PF:750-761) and the 1/(Lz*Lx) scaling (PF:1176-1177).
    """
    K = r.shape[0]
    h = p.Mz // 2
    x = np.fft.fft(r, axis=2)[..., :p.Mx]
    g = np.fft.fft(x, axis=1)
    f = np.empty((K, p.Mz, p.Mx), dtype=np.complex128)
    f[:, :h + 1] = g[:, :h + 1]
    if h > 1:
        f[:, h + 1:] = g[:, p.Lz - (h - 1):]
    return f / float(p.Lz * p.Lx)


# --------------------------------------------------------------------------
# wall-normal padded product (getlamb PF:1657-1706, getlamb1 PF:1758-1809)
# --------------------------------------------------------------------------


def _even_ext(t: np.ndarray, m2: int) -> np.ndarray:
    """in(1..2*m2) of PF:1672-1677: t_0..t_m2 then t_{m2-1}..t_1 along axis 0."""
    return np.concatenate([t, t[m2 - 1:0:-1]], axis=0)


def cheb_padded_product(a: np.ndarray, b: np.ndarray, n: int, shift: float = 0.0) -> np.ndarray:
    """Degree-N truncated Chebyshev product of two coefficient columns (axis 0).

    Pads both to degree 2N (PF:1659-1667), evaluates the packed pair a+ib on
    the 2N+1 Gauss-Lobatto points with one length-4N FFT (plan13, PF:1672-1680),
    forms Re*Im/4 (PF:1682-1684), subtracts ``shift`` (getlamb1's re_p,
    PF:1787), analyses with a second length-4N FFT (PF:1686-1700) and keeps
    degrees 0..N (PF:1702-1704).
    """
    m2 = 2 * n
    pad = np.zeros((m2 + 1,) + a.shape[1:], dtype=np.float64)
    ta = pad.copy()
    tb = pad.copy()
    ta[:n + 1] = a
    tb[:n + 1] = b
    packed = ta + 1j * tb
    ext = _even_ext(packed, m2)
    ext[0] *= 2.0
    ext[m2] *= 2.0
    out = np.fft.fft(ext, axis=0)[:m2 + 1]
    prod = out.real * out.imag / 4.0
    if shift:
        prod = prod - shift
    ext2 = _even_ext(prod.astype(np.complex128), m2)
    res = np.fft.fft(ext2, axis=0).real[:m2 + 1] / float(m2)
    res[0] /= 2.0
    res[m2] /= 2.0
    return res[:n + 1]


def getlamb(u: np.ndarray, omega: np.ndarray, p: Params, shift: float = 0.0) -> np.ndarray:
    """Exactly dealiased spectral product u*omega (getlamb PF:1625-1712).

    ``shift`` = re_p gives getlamb1 (PF:1726-1815).
    """
    up = fourier_to_phys(u, p)
    op = fourier_to_phys(omega, p)
    prod = cheb_padded_product(up, op, p.N, shift)
    return phys_to_fourier(prod, p)


def points_from_coeffs(a: np.ndarray, n: int) -> np.ndarray:
    """Values at the N+1 points y_j = cos(pi j/N) (plan12, PF:1573-1609)."""
    ext = np.concatenate([a, a[n - 1:0:-1]], axis=0).astype(np.complex128)
    ext[0] *= 2.0
    ext[n] *= 2.0
    return np.fft.fft(ext, axis=0)[:n + 1].real / 2.0


# --------------------------------------------------------------------------
# state (PC:28-38) and the nonlinear step (getnonlinear PF:1380-1612)
# --------------------------------------------------------------------------

FIELD_NAMES = ("vnp1", "omegaynp1", "rhsvn", "rhsomegayn", "unp1", "wnp1",
               "omegaxnp1", "omegaznp1")
MEAN_NAMES = ("clamb1n", "clamb1nm1", "clamb3n", "clamb3nm1", "au0", "w0")


def zero_state(p: Params) -> dict:
    st = {k: np.zeros((p.Ny, p.Mz, p.Mx), dtype=np.complex128) for k in FIELD_NAMES}
    st.update({k: np.zeros(p.Ny, dtype=np.complex128) for k in MEAN_NAMES})
    return st


def copy_state(st: dict) -> dict:
    return {k: v.copy() for k, v in st.items()}


def getnonlinear(st: dict, p: Params) -> None:
    """getnonlinear (PF:1380-1612), in place on the reference's arrays."""
    alpha, beta = wavenumbers(p)
    u, v, w = st["unp1"], st["vnp1"], st["wnp1"]
    ox, oy, oz = st["omegaxnp1"], st["omegaynp1"], st["omegaznp1"]
    # six products, lamb vector = omega x u - re_p x^ (PF:1438-1473)
    lamb1 = getlamb(oy, w, p, shift=p.re_p) - getlamb(oz, v, p)
    lamb2 = getlamb(oz, u, p) - getlamb(ox, w, p)
    lamb3 = getlamb(ox, v, p) - getlamb(oy, u, p)
    # mean-mode lamb saved into the "nm1" slot then swapped (PF:1482-1487, 1554-1570)
    new_cl1 = lamb1[:, 0, 0].copy()
    new_cl3 = lamb3[:, 0, 0].copy()
    # h_v = k^2 L_y + i alpha D L_x + i beta D L_z ; h_g = i alpha L_z - i beta L_x (PF:1502-1540)
    k2 = alpha[None, None, :] ** 2 + beta[None, :, None] ** 2
    a3 = alpha[None, None, :]
    b3 = beta[None, :, None]
    dlx = dz(lamb1)
    dlz = dz(lamb3)
    hv = k2 * lamb2 + 1j * a3 * dlx
    hv = hv + 1j * b3 * dlz
    hg = 1j * a3 * lamb3 - 1j * b3 * lamb1
    # the reference leaves L_x in omegaxnp1 and L_z in omegaznp1 (PF:1463-1470, 1492)
    st["omegaxnp1"] = lamb1
    st["omegaznp1"] = lamb3
    # history swap: unp1,wnp1 <- h^{n-1}; rhsvn,rhsomegayn <- h^n (PF:1542-1553)
    st["unp1"] = st["rhsvn"]
    st["wnp1"] = st["rhsomegayn"]
    st["clamb1nm1"] = st["clamb1n"]
    st["clamb3nm1"] = st["clamb3n"]
    # h^n and the mean lamb -> values at the N+1 points (plan12, PF:1573-1609)
    st["rhsvn"] = _points_complex(hv, p.N)
    st["rhsomegayn"] = _points_complex(hg, p.N)
    st["clamb1n"] = _points_complex(new_cl1, p.N)
    st["clamb3n"] = _points_complex(new_cl3, p.N)


def _points_complex(a: np.ndarray, n: int) -> np.ndarray:
    return points_from_coeffs(a.real, n) + 1j * points_from_coeffs(a.imag, n)


# --------------------------------------------------------------------------
# operators (get_coef_v PF:3348-3381, get_coef_omegay PF:3432-3455,
# get_coef_u0 PF:3493-3511) and right-hand sides (PF:3385-3425, 3458-3489,
# 3515-3534)
# --------------------------------------------------------------------------


class Operators:
    """T matrices plus per-k^2 operator construction."""

    def __init__(self, p: Params):
        self.p = p
        self.t0, self.t1, self.t2, self.t4 = scheb(p.N)

    def coef_v(self, kmn: float) -> np.ndarray:
        """get_coef_v (PF:3348-3381): rows 4..N at y_{i-2}, rows 0..3 wall BCs."""
        p, n = self.p, self.p.N
        t0, t2, t4 = self.t0, self.t2, self.t4
        a = np.zeros((n + 1, n + 1))
        rows = np.arange(2, n - 1)
        a[4:] = (-t4[rows] * p.dt * 0.5 / p.re + (1 + kmn * p.dt / p.re) * t2[rows]
                 - kmn * (1 + kmn * p.dt * 0.5 / p.re) * t0[rows])
        j = np.arange(n + 1)
        sgn = np.where(j % 2 == 0, 1.0, -1.0)
        a[0] = 1.0
        a[1] = sgn
        a[2] = j.astype(np.float64) ** 2
        a[3] = -sgn * j.astype(np.float64) ** 2
        return a

    def rhs_op_v(self, kmn: float) -> np.ndarray:
        """Explicit CN operator of get_coef_rhs (PF:3408-3409), rows = points 2..N-2."""
        p, n = self.p, self.p.N
        rows = np.arange(2, n - 1)
        return (self.t4[rows] * p.dt * 0.5 / p.re + (1 - kmn * p.dt / p.re) * self.t2[rows]
                - kmn * (1 - kmn * p.dt * 0.5 / p.re) * self.t0[rows])

    def coef_omegay(self, kmn: float) -> np.ndarray:
        """get_coef_omegay (PF:3432-3455): rows 2..N at y_{i-1}, rows 0,1 Dirichlet."""
        p, n = self.p, self.p.N
        a = np.zeros((n + 1, n + 1))
        rows = np.arange(1, n)
        a[2:] = -self.t2[rows] * p.dt * 0.5 / p.re + (1 + kmn * p.dt * 0.5 / p.re) * self.t0[rows]
        j = np.arange(n + 1)
        a[0] = 1.0
        a[1] = np.where(j % 2 == 0, 1.0, -1.0)
        return a

    def rhs_op_omegay(self, kmn: float) -> np.ndarray:
        """Explicit operator of get_rhs_omegay (PF:3475), rows = points 1..N-1."""
        p, n = self.p, self.p.N
        rows = np.arange(1, n)
        return self.t2[rows] * p.dt * 0.5 / p.re + (1 - kmn * p.dt * 0.5 / p.re) * self.t0[rows]

    def coef_u0(self) -> np.ndarray:
        """get_coef_u0 (PF:3493-3511)."""
        p, n = self.p, self.p.N
        a = np.zeros((n + 1, n + 1))
        rows = np.arange(1, n)
        a[2:] = -self.t2[rows] * p.dt * 0.5 / p.re + self.t0[rows]
        j = np.arange(n + 1)
        a[0] = 1.0
        a[1] = np.where(j % 2 == 0, 1.0, -1.0)
        return a

    def rhs_op_u0(self) -> np.ndarray:
        """get_rhs_u0 (PF:3526), rows = points 1..N-1."""
        p, n = self.p, self.p.N
        rows = np.arange(1, n)
        return self.t0[rows] + p.dt * 0.5 / p.re * self.t2[rows]


def gsij(aa: np.ndarray, af: np.ndarray) -> np.ndarray:
    """Full-pivot Gaussian elimination of PF:3540-3618 (ka=0 branch).

    The pivot is the first entry of largest modulus in (row, column) scan order
    among unused rows and columns (PF:3557-3568); elimination updates unused
    rows/columns only (PF:3576-3585); back substitution follows PF:3603-3613.
    With ka=1 the reference replays the same pivot order on the same matrix,
    which is the same arithmetic.
    """
    aa = np.array(aa, dtype=np.complex128, copy=True)
    af = np.array(af, dtype=np.complex128, copy=True)
    n = aa.shape[0]
    ki = np.ones(n, dtype=bool)
    kj = np.ones(n, dtype=bool)
    ik, jk = [], []
    im = jm = 0
    for k in range(n):
        mag = np.sqrt(aa.real ** 2 + aa.imag ** 2)
        mag = np.where(ki[:, None] & kj[None, :], mag, -1.0)
        flat = int(np.argmax(mag))
        im, jm = divmod(flat, n)
        ik.append(im)
        jk.append(jm)
        kj[jm] = False
        ki[im] = False
        if k == n - 1:
            break
        add = 1.0 / aa[im, jm]
        rows = np.nonzero(ki)[0]
        cols = np.nonzero(kj)[0]
        fac = aa[rows, jm] * add
        aa[np.ix_(rows, cols)] -= fac[:, None] * aa[im, cols][None, :]
        af[rows] -= fac * af[im]
    aa[im, jm] = af[im] / aa[im, jm]
    for k in range(n - 2, -1, -1):
        i, j = ik[k], jk[k]
        acc = af[i]
        for lj in range(k + 1, n):
            jj = jk[lj]
            acc = acc - aa[i, jj] * aa[im, jj]
        aa[im, j] = acc / aa[i, j]
    return aa[im, :].copy()


# --------------------------------------------------------------------------
# per-mode solves (PF:242-394), kinematics (PF:2347-2407), mean mode (PF:401-474)
# --------------------------------------------------------------------------


class Stepper:
    """Time loop of PF:220-532 from a pre-solve state: [solve ; getnonlinear]."""

    def __init__(self, p: Params):
        self.p = p
        self.ops = Operators(p)
        self._inv_cache = {}
        n = p.N
        self._u0_mat = self.ops.coef_u0()
        self._u0_rhs = self.ops.rhs_op_u0()
        self.solve_modes = []
        alpha, beta = wavenumbers(p)
        for i in range(p.Mx):
            for j in range(p.Mz):
                kmn = alpha[i] ** 2 + beta[j] ** 2
                if j == p.Mz // 2 or kmn == 0.0:       # PF:247
                    continue
                self.solve_modes.append((i, j, kmn))
        self._n = n

    def _inverses(self, kmn: float):
        """coef_a / coef_e inverses (zgetrf+zgetri, PF:181-204), real-valued."""
        hit = self._inv_cache.get(kmn)
        if hit is None:
            av = np.linalg.inv(self.ops.coef_v(kmn))
            ae = np.linalg.inv(self.ops.coef_omegay(kmn))
            hit = (av, ae, self.ops.rhs_op_v(kmn), self.ops.rhs_op_omegay(kmn))
            self._inv_cache[kmn] = hit
        return hit

    def solve(self, st: dict, chunk: int = 256) -> None:
        """Per-mode solve loop (PF:242-394) batched over modes.

        Each mode's arithmetic is the reference's: rhs = explicit operator
        times v^n (and omega_y^n) plus the AB2 forcing at the collocation
        points (PF:252-257, 3396-3423; PF:335-340, 3472-3487), x = A^-1 rhs
        (a_multiply_x PF:4123-4139), then the kinematics of PF:2357-2403.
        """
        p, n, dt = self.p, self.p.N, self.p.dt
        alpha, beta = wavenumbers(p)
        v, oy = st["vnp1"], st["omegaynp1"]
        u, w = st["unp1"], st["wnp1"]
        ox, oz = st["omegaxnp1"], st["omegaznp1"]
        hv, hg = st["rhsvn"], st["rhsomegayn"]
        modes = self.solve_modes
        for c0 in range(0, len(modes), chunk):
            blk = modes[c0:c0 + chunk]
            ii = np.array([m[0] for m in blk])
            jj = np.array([m[1] for m in blk])
            kk = np.array([m[2] for m in blk])
            mats = [self._inverses(k) for k in kk]
            av = np.stack([m[0] for m in mats])
            ae = np.stack([m[1] for m in mats])
            bv = np.stack([m[2] for m in mats])
            be = np.stack([m[3] for m in mats])
            # v equation (PF:252-331)
            vm = v[:, jj, ii].T                                       # (nc, Ny)
            f = (3.0 * dt / 2.0 * hv[:, jj, ii] - dt / 2.0 * u[:, jj, ii]).T
            rhs = np.zeros((len(blk), n + 1), dtype=np.complex128)
            rhs[:, 4:] = np.einsum("mpl,ml->mp", bv, vm) + f[:, 2:n - 1]
            vn = np.einsum("mkl,ml->mk", av, rhs)
            # omega_y equation (PF:335-384)
            om = oy[:, jj, ii].T
            f = (3.0 * dt / 2.0 * hg[:, jj, ii] - dt / 2.0 * w[:, jj, ii]).T
            rhs = np.zeros((len(blk), n + 1), dtype=np.complex128)
            rhs[:, 2:] = np.einsum("mpl,ml->mp", be, om) + f[:, 1:n]
            on = np.einsum("mkl,ml->mk", ae, rhs)
            # kinematics (getvelocityandvorticity PF:2357-2403)
            a_ = alpha[ii][None, :]
            b_ = beta[jj][None, :]
            k2 = kk[None, :]
            vt, ot = vn.T, on.T                                         # (Ny, nc)
            dv = dz(vt)
            uu = (1j * a_ * dv - 1j * b_ * ot) / k2
            ww = (1j * b_ * dv + 1j * a_ * ot) / k2
            v[:, jj, ii] = vt
            oy[:, jj, ii] = ot
            u[:, jj, ii] = uu
            w[:, jj, ii] = ww
            ox[:, jj, ii] = dz(ww) - 1j * b_ * vt
            oz[:, jj, ii] = -dz(uu) + 1j * a_ * vt
        self.solve_mean(st)

    def solve_mean(self, st: dict) -> None:
        """Mean mode U(y), W(y) (PF:401-474) with gsij."""
        p, n, dt = self.p, self.p.N, self.p.dt
        for cn, cnm1, mean, vel, vort, sgn in (("clamb1n", "clamb1nm1", "au0", "unp1", "omegaznp1", -1.0),
                                               ("clamb3n", "clamb3nm1", "w0", "wnp1", "omegaxnp1", 1.0)):
            f = -1.0 * dt / 2.0 * (3.0 * st[cn] - st[cnm1])
            rhs = np.zeros(n + 1, dtype=np.complex128)
            rhs[2:] = self._u0_rhs @ st[mean] + f[1:n]
            x = gsij(self._u0_mat, rhs)
            st[mean] = x
            st[vel][:, 0, 0] = x
            st[vort][:, 0, 0] = sgn * dz(x)

    def step(self, st: dict, nsteps: int = 1) -> None:
        """nsteps x [per-mode solve ; mean mode ; getnonlinear] from a pre-solve state."""
        for _ in range(nsteps):
            self.solve(st)
            getnonlinear(st, self.p)


def kinematics_all(st: dict, p: Params) -> None:
    """getvelocityandvorticity for every k^2 != 0 mode (PF:2288-2295 in input),
    plus the mean-mode velocity/vorticity from au0, w0 (PF:423-472)."""
    alpha, beta = wavenumbers(p)
    k2 = alpha[None, None, :] ** 2 + beta[None, :, None] ** 2
    mask = np.ones((1, p.Mz, p.Mx), dtype=bool)
    mask[0, p.Mz // 2, :] = False          # Nyquist column (PF:247)
    mask[0, 0, 0] = False                  # k^2 = 0 (PF:2365)
    k2s = np.where(mask, k2, 1.0)
    a_ = alpha[None, None, :]
    b_ = beta[None, :, None]
    vn, on = st["vnp1"], st["omegaynp1"]
    dv = dz(vn)
    uu = (1j * a_ * dv - 1j * b_ * on) / k2s
    ww = (1j * b_ * dv + 1j * a_ * on) / k2s
    ox = dz(ww) - 1j * b_ * vn
    oz = -dz(uu) + 1j * a_ * vn
    for name, val in (("unp1", uu), ("wnp1", ww), ("omegaxnp1", ox), ("omegaznp1", oz)):
        st[name] = np.where(mask, val, st[name])
    st["unp1"][:, 0, 0] = st["au0"]
    st["wnp1"][:, 0, 0] = st["w0"]
    st["omegaznp1"][:, 0, 0] = -dz(st["au0"])
    st["omegaxnp1"][:, 0, 0] = dz(st["w0"])


def initial_presolve_state(v: np.ndarray, oy: np.ndarray, au0: np.ndarray, w0: np.ndarray,
                           p: Params) -> dict:
    """Cold start (initial, PF:1932-1946): fields from (v, omega_y, U, W), h^{-1}=0,
    then one getnonlinear gives the pre-solve state the time loop starts from."""
    st = zero_state(p)
    st["vnp1"] = np.array(v, dtype=np.complex128, copy=True)
    st["omegaynp1"] = np.array(oy, dtype=np.complex128, copy=True)
    st["au0"] = np.array(au0, dtype=np.complex128, copy=True)
    st["w0"] = np.array(w0, dtype=np.complex128, copy=True)
    kinematics_all(st, p)
    getnonlinear(st, p)
    return st


# --------------------------------------------------------------------------
# Unpadded physical output and plane statistics (test infrastructure)


def physical_field(F, p: Params):
    """transform(1, .) (PF:948-961, fourtrx_zcomp(+1) PF:774-860) of one
    spectral field F = complex[Ny][Mz][Mx] (coefficients): values at
    y_k = cos(pi k/N) (plain Chebyshev sum, PF:951-957), span inverse with the
    Nyquist slot zeroed (PF:792-794), streamwise Hermitian fill and Re
    (getall PF:1350-1362, PF:806-818).  Returns real[Ny][nny][nnx]."""
    F = np.asarray(F, dtype=np.complex128)
    Ny, Mz, Mx = F.shape
    N = Ny - 1
    k = np.arange(Ny)
    T = np.cos(np.pi * np.outer(k, k) / N)
    pts = np.einsum("kl,lzx->kzx", T, F)
    pts[:, Mz // 2, :] = 0.0
    S = np.fft.ifft(pts, axis=1) * Mz                 # e^{+i}, unnormalised
    nnx = 2 * Mx
    G = np.zeros((Ny, Mz, nnx), dtype=np.complex128)
    G[:, :, 0] = S[:, :, 0].real
    G[:, :, 1:Mx] = S[:, :, 1:Mx]
    G[:, :, nnx - Mx + 1:] = np.conj(S[:, :, 1:Mx][:, :, ::-1])
    return (np.fft.ifft(G, axis=2) * nnx).real


def plane_statistics(u, v, w):
    """Plane means, covariances about the plane mean and extrema of physical
    u, v, w (statistics, PF:2825-2868 and PF:2869-3020): real[12][Ny], rows
    ua va wa uu vv ww uv uw vw max|u| max|v| max|w|."""
    out = []
    m = [f.mean(axis=(1, 2)) for f in (u, v, w)]
    d = [f - mm[:, None, None] for f, mm in zip((u, v, w), m)]
    out += m
    out += [(d[0] * d[0]).mean(axis=(1, 2)), (d[1] * d[1]).mean(axis=(1, 2)), (d[2] * d[2]).mean(axis=(1, 2)),
            (d[0] * d[1]).mean(axis=(1, 2)), (d[0] * d[2]).mean(axis=(1, 2)), (d[1] * d[2]).mean(axis=(1, 2))]
    out += [np.abs(f).max(axis=(1, 2)) for f in (u, v, w)]
    return np.array(out)
