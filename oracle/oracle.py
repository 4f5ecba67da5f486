"""ctypes wrapper of the FP64 CPU oracle (oracle/spns_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, where it is the checker.  The product
(spns_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libspns_oracle.so")
_lib = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
        ("Mx", ctypes.c_int), ("My", ctypes.c_int), ("Mz", ctypes.c_int),
        ("nu", ctypes.c_double), ("C_Smag", ctypes.c_double), ("C_At", ctypes.c_double),
        ("Lx", ctypes.c_int), ("Ly", ctypes.c_int), ("Lz", ctypes.c_int),
        ("omega_x", ctypes.c_double), ("omega_y", ctypes.c_double), ("omega_z", ctypes.c_double),
        ("isLES", ctypes.c_int), ("isRotating", ctypes.c_int), ("isNotCubic", ctypes.c_int),
        ("isSelfAdaptive", ctypes.c_int), ("isCCC", ctypes.c_int),
        ("delta", ctypes.c_double), ("fixed_dt", ctypes.c_double),
        ("force_on", ctypes.c_int), ("force_eps", ctypes.c_double),
        ("force_kmin", ctypes.c_int), ("force_kmax", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [P]
        for name in ["orc_antitransform", "orc_step", "orc_recalc_timestep", "orc_ccc"]:
            getattr(L, name).argtypes = [P]
        L.orc_steps.argtypes = [P, ctypes.c_int]
        for name in ["orc_energy", "orc_enstrophy", "orc_relambda"]:
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = ctypes.c_double
        L.orc_energy_cascade.argtypes = [P, D]
        L.orc_energy_cascade.restype = ctypes.c_double
        L.orc_forcing_relambda.argtypes = [P, ctypes.c_double]
        L.orc_forcing_energy_cascade.argtypes = [P, D, ctypes.c_long, ctypes.c_int]
        L.orc_init_k41.argtypes = [P, ctypes.c_int]
        L.orc_init_random.argtypes = [P, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        L.orc_splitmix.argtypes = [ctypes.c_uint64]
        L.orc_splitmix.restype = ctypes.c_uint64
        L.orc_rand_unif.argtypes = [ctypes.c_uint64] + [ctypes.c_int] * 4
        L.orc_rand_unif.restype = ctypes.c_double
        L.orc_init_from_physical.argtypes = [P, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D]
        L.orc_finish_init.argtypes = [P, ctypes.c_int]
        L.orc_nmodes.argtypes = [P]
        L.orc_nmodes.restype = ctypes.c_long
        L.orc_get_state.argtypes = [P, D]
        L.orc_set_state.argtypes = [P, D]
        L.orc_get_rhist.argtypes = [P, D]
        L.orc_velocity_physical.argtypes = [P, D]
        L.orc_parseval.argtypes = [P, D]
        L.orc_invariants_point.argtypes = [D, D]
        L.orc_keep_start.argtypes = [P]
        L.orc_invariant_spectra.argtypes = [P, D]
        L.orc_get_scalars.argtypes = [P, D]
        L.orc_set_nu.argtypes = [P, ctypes.c_double]
        L.orc_last_rad_max.argtypes = [P]
        L.orc_last_rad.argtypes = [P]
        L.orc_fft_c2r.argtypes = [ctypes.c_int] * 3 + [D, D]
        L.orc_fft_r2c.argtypes = [ctypes.c_int] * 3 + [D, D]
        L.orc_fft1d.argtypes = [ctypes.c_int, D, ctypes.c_int]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_prefault.argtypes = [P]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


SCALAR_NAMES = ["time", "dt", "nu", "maxvel", "ek", "om", "relambda", "skew_s2", "skew_s3", "iter", "kappa"]


class Oracle:
    """One solver instance.  Keyword arguments follow the HIT constructor
    (src/HIT_LES.h:23-29) plus the extensions of orc_params."""

    def __init__(self, Nx, Ny=None, Nz=None, Mx=None, My=None, Mz=None, nu=1e-2, C_Smag=0.0, C_At=1.0,
                 Lx=1, Ly=1, Lz=1, omega=(0.0, 0.0, 0.0), isNotCubic=False, isSelfAdaptive=False,
                 isCCC=False, delta=0.0, fixed_dt=0.0, force_eps=None, force_kmin=1, force_kmax=2):
        Ny = Nx if Ny is None else Ny
        Nz = Nx if Nz is None else Nz
        Mx = (3 * Nx) // 2 if Mx is None else Mx
        My = (3 * Ny) // 2 if My is None else My
        Mz = (3 * Nz) // 2 if Mz is None else Mz
        p = OrcParams()
        p.Nx, p.Ny, p.Nz, p.Mx, p.My, p.Mz = Nx, Ny, Nz, Mx, My, Mz
        p.nu, p.C_Smag, p.C_At = nu, C_Smag, C_At
        p.Lx, p.Ly, p.Lz = Lx, Ly, Lz
        p.omega_x, p.omega_y, p.omega_z = omega
        p.isLES = int(C_Smag > 0)
        p.isRotating = int(any(abs(w) > 0 for w in omega))
        p.isNotCubic = int(isNotCubic)
        p.isSelfAdaptive = int(isSelfAdaptive)
        p.isCCC = int(isCCC)
        p.delta, p.fixed_dt = delta, fixed_dt
        p.force_on = int(force_eps is not None)
        p.force_eps = 0.0 if force_eps is None else force_eps
        p.force_kmin, p.force_kmax = force_kmin, force_kmax
        self.p = p
        self.L = lib()
        self.h = self.L.orc_create(ctypes.byref(p))
        if not self.h:
            raise ValueError("oracle: unsupported grid")
        self.nmodes = self.L.orc_nmodes(self.h)
        self.last_rad_max = self.L.orc_last_rad_max(self.h)
        self.last_rad = self.L.orc_last_rad(self.h)
        self.shape = (Mx, My, Mz)
        self.nk = (2 * (Ny // 2) + 1, 2 * (Nx // 2) + 1, Nz // 2 + 1)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    # --- initial fields ---
    def set_state(self, state):
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(3, self.nmodes)
        self.L.orc_set_state(self.h, _dp(s.view(np.float64)))

    def get_state(self):
        out = np.zeros((3, self.nmodes), dtype=np.complex128)
        self.L.orc_get_state(self.h, _dp(out.view(np.float64)))
        return out

    def get_rhist(self):
        out = np.zeros((3, self.nmodes), dtype=np.complex128)
        self.L.orc_get_rhist(self.h, _dp(out.view(np.float64)))
        return out

    def init_k41(self, nranks=4):
        self.L.orc_init_k41(self.h, nranks)

    def init_random(self, seed=0, k0=4.0, urms=1.0):
        """Seeded random-phase solenoidal field, E(k) ~ k^4 exp(-2 (k/k0)^2),
        E_tot = 1.5 u_rms^2 (the GPU's spns_init_random restated)."""
        self.L.orc_init_random(self.h, seed, k0, urms)

    def init_from_physical(self, u, v, w):
        nf = u.shape
        u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
        ekf, ensf = ctypes.c_double(), ctypes.c_double()
        self.L.orc_init_from_physical(self.h, _dp(u), _dp(v), _dp(w), nf[0], nf[1], nf[2],
                                      ctypes.byref(ekf), ctypes.byref(ensf))
        return ekf.value, ensf.value

    def finish_init(self, isReLambda=False):
        self.L.orc_finish_init(self.h, int(isReLambda))

    def prefault(self):
        """Touch every page of the field arrays (values unchanged), so a
        timed step does not pay first-touch page faults."""
        self.L.orc_prefault(self.h)

    # --- stepping / forcing ---
    def step(self, n=1):
        self.L.orc_steps(self.h, n)

    def recalc_timestep(self):
        self.L.orc_recalc_timestep(self.h)

    def antitransform(self):
        self.L.orc_antitransform(self.h)

    def set_nu(self, nu):
        self.L.orc_set_nu(self.h, nu)

    def forcing_relambda(self, relambda):
        self.L.orc_forcing_relambda(self.h, relambda)

    def forcing_energy_cascade(self, ek, last_input_rad, nullify=False):
        ek = np.ascontiguousarray(ek, dtype=np.float64)
        self.L.orc_forcing_energy_cascade(self.h, _dp(ek), last_input_rad, int(nullify))

    def ccc(self):
        self.L.orc_ccc(self.h)

    # --- diagnostics ---
    def energy(self):
        return self.L.orc_energy(self.h)

    def enstrophy(self):
        return self.L.orc_enstrophy(self.h)

    def relambda(self):
        return self.L.orc_relambda(self.h)

    def spectrum(self):
        ek = np.zeros(self.last_rad_max + 1)
        self.L.orc_energy_cascade(self.h, _dp(ek))
        return ek

    def parseval(self):
        out = np.zeros(3)
        self.L.orc_parseval(self.h, _dp(out))
        return out

    def keep_start(self):
        """Keep the state at the start of every step (for invariant_spectra)."""
        self.L.orc_keep_start(self.h)

    def invariant_spectra(self):
        """Shell spectra of QG, QS, RG, RS, V2 of the velocity gradients at the
        start of the last step: array [5][last_rad_max + 1]."""
        out = np.zeros((5, self.last_rad_max + 1))
        self.L.orc_invariant_spectra(self.h, _dp(out))
        return out

    def velocity_physical(self):
        Mx, My, Mz = self.shape
        out = np.zeros((3, Mx, My, Mz))
        self.L.orc_velocity_physical(self.h, _dp(out))
        return out

    def scalars(self):
        s = np.zeros(len(SCALAR_NAMES))
        self.L.orc_get_scalars(self.h, _dp(s))
        return dict(zip(SCALAR_NAMES, s.tolist()))


def invariants_point(g):
    """QG, QS, RG, RS, V2 of one velocity-gradient tensor g[i][j] = d u_i / d x_j."""
    g = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(9))
    out = np.zeros(5)
    lib().orc_invariants_point(_dp(g), _dp(out))
    return out


def fft_c2r(cin, Mx, My, Mz):
    """cin: complex [My][Mx][Mz//2+1] -> real [Mx][My][2(Mz//2+1)] (unnormalised)."""
    cin = np.ascontiguousarray(cin, dtype=np.complex128)
    out = np.zeros((Mx, My, 2 * (Mz // 2 + 1)))
    lib().orc_fft_c2r(Mx, My, Mz, _dp(cin.view(np.float64)), _dp(out))
    return out


def fft_r2c(rin, Mx, My, Mz):
    """rin: real [Mx][My][2(Mz//2+1)] -> complex [My][Mx][Mz//2+1] (unnormalised)."""
    rin = np.ascontiguousarray(rin, dtype=np.float64)
    out = np.zeros((My, Mx, Mz // 2 + 1), dtype=np.complex128)
    lib().orc_fft_r2c(Mx, My, Mz, _dp(rin), _dp(out.view(np.float64)))
    return out


def fft1d(x, sign=-1):
    x = np.array(x, dtype=np.complex128)
    lib().orc_fft1d(len(x), _dp(x.view(np.float64)), sign)
    return x
