"""MI355X-native pseudospectral HIT step (drop-in for SpNS's HIT time step).

Python binding of the C ABI in include/spns.h (libspns.so, built in-tree by
`python -m spns_amd.build`).  There is no CPU fallback: if the HIP library
is missing or no GPU is present, constructing a Solver raises.

    s = Solver(Nx=32, nu=0.01, fixed_dt=1e-3)
    s.init_random(seed=1, k0=4.0)
    s.finish_init()
    s.step(200)
    print(s.scalars(), s.spectrum())
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspns.so")
PREC_FP64 = 0
PREC_FP32_SPLIT = 1
# return codes of the C ABI (include/spns.h)
SPNS_OK, SPNS_EINVAL, SPNS_EUNSUPPORTED, SPNS_ENOMEM, SPNS_EHIP, SPNS_ENONFINITE, SPNS_ECHECK = range(7)
_lib = None


class Config(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32), ("Nz", ctypes.c_int32),
        ("Mx", ctypes.c_int32), ("My", ctypes.c_int32), ("Mz", ctypes.c_int32),
        ("nu", ctypes.c_double), ("C_Smag", ctypes.c_double), ("C_At", ctypes.c_double),
        ("Lx", ctypes.c_int32), ("Ly", ctypes.c_int32), ("Lz", ctypes.c_int32),
        ("omega_x", ctypes.c_double), ("omega_y", ctypes.c_double), ("omega_z", ctypes.c_double),
        ("isNotCubic", ctypes.c_int32), ("isSelfAdaptive", ctypes.c_int32),
        ("isComplexConjugateCorrected", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("delta", ctypes.c_double), ("fixed_dt", ctypes.c_double),
        ("force_on", ctypes.c_int32), ("force_eps", ctypes.c_double),
        ("force_kmin", ctypes.c_int32), ("force_kmax", ctypes.c_int32),
        ("device", ctypes.c_int32), ("use_graph", ctypes.c_int32),
        ("zpass_variant", ctypes.c_int32),
        ("gradient_invariants", ctypes.c_int32),
        ("keep_step_start", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


class Scalars(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double), ("nu", ctypes.c_double),
                ("maxvel", ctypes.c_double), ("ek", ctypes.c_double), ("om", ctypes.c_double),
                ("relambda", ctypes.c_double), ("skewness", ctypes.c_double), ("kappa", ctypes.c_double),
                ("iter", ctypes.c_int64)]


EXPORTS = [
    "spns_create", "spns_reconfigure", "spns_destroy", "spns_device_bytes", "spns_nmodes", "spns_nbins", "spns_last_rad", "spns_last_error",
    "spns_set_state", "spns_get_state", "spns_get_state_planes", "spns_get_rhist", "spns_init_from_physical", "spns_init_k41",
    "spns_init_random", "spns_finish_init", "spns_step", "spns_recalculate_timestep", "spns_set_nu",
    "spns_forcing_reynolds_lambda", "spns_recalculate_reynolds_lambda", "spns_forcing_energy_cascade",
    "spns_complex_conjugate_correction", "spns_energy_cascade", "spns_pass_kernels", "spns_get_scalars",
    "spns_velocity_physical", "spns_velocity_physical_step_start", "spns_check_phys_four_energy",
    "spns_recalculate_velocity_antitransform", "spns_synchronize", "spns_query", "spns_device_memory", "spns_copy_bandwidth", "spns_last_step_ms", "spns_invariant_spectra",
    "spns_invariant_spectra_noalloc",
    "spns_profile_passes",
    "spns_fft_create", "spns_fft_destroy", "spns_fft_r2c", "spns_fft_c2r", "spns_fft_normalize",
    "spns_fft_synchronize", "spns_fft_last_error",
    "spns_batch_create", "spns_batch_destroy", "spns_batch_attach", "spns_batch_detach", "spns_batch_step",
    "spns_batch_query", "spns_batch_synchronize", "spns_batch_slot_scalars", "spns_batch_fetch", "spns_batch_last_error",
    "spns_host_alloc", "spns_host_free",
]


def load_library(path=LIB_PATH):
    """Loads libspns.so and declares the C ABI.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("spns_amd: %s is not built; run `python -m spns_amd.build`" % path)
    L = ctypes.CDLL(path)
    P, D, I32, I64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32, ctypes.c_int64
    sig = {
        "spns_create": (I32, [ctypes.POINTER(Config), ctypes.POINTER(P)]),
        "spns_destroy": (I32, [P]),
        "spns_reconfigure": (I32, [P, ctypes.POINTER(Config)]),
        "spns_device_bytes": (I64, [ctypes.POINTER(Config)]),
        "spns_nmodes": (I64, [P]),
        "spns_nbins": (I32, [P]),
        "spns_last_rad": (I32, [P]),
        "spns_last_error": (ctypes.c_char_p, [P]),
        "spns_set_state": (I32, [P, D]),
        "spns_get_state": (I32, [P, D]),
        "spns_get_rhist": (I32, [P, D]),
        "spns_init_from_physical": (I32, [P, D, D, D, I32, I32, I32, D, D]),
        "spns_init_k41": (I32, [P, I32]),
        "spns_init_random": (I32, [P, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
        "spns_finish_init": (I32, [P, I32]),
        "spns_step": (I32, [P, I32]),
        "spns_recalculate_timestep": (I32, [P]),
        "spns_set_nu": (I32, [P, ctypes.c_double]),
        "spns_forcing_reynolds_lambda": (I32, [P, ctypes.c_double]),
        "spns_recalculate_reynolds_lambda": (I32, [P, D]),
        "spns_forcing_energy_cascade": (I32, [P, D, I64, I32]),
        "spns_complex_conjugate_correction": (I32, [P]),
        "spns_energy_cascade": (I32, [P, D, I32, D]),
        "spns_get_scalars": (I32, [P, ctypes.POINTER(Scalars)]),
        "spns_pass_kernels": (I32, [P, ctypes.POINTER(I32), ctypes.POINTER(I32)]),
        "spns_velocity_physical": (I32, [P, D]),
        "spns_velocity_physical_step_start": (I32, [P, D]),
        "spns_recalculate_velocity_antitransform": (I32, [P]),
        "spns_check_phys_four_energy": (I32, [P, ctypes.c_double, D, D, D]),
        "spns_synchronize": (I32, [P]),
        "spns_query": (I32, [P, ctypes.POINTER(ctypes.c_int32)]),
        "spns_device_memory": (I32, [I32, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "spns_copy_bandwidth": (I32, [I32, I64, I32, D]),
        "spns_get_state_planes": (I32, [P, I32, I32, D]),
        "spns_invariant_spectra": (I32, [P, I32, D]),
        "spns_invariant_spectra_noalloc": (I32, [P, I32, D]),
        "spns_last_step_ms": (I32, [P, D]),
        "spns_profile_passes": (I32, [P, I32, D]),
        "spns_fft_create": (I32, [I32, I32, I32, I32, I32, ctypes.POINTER(P)]),
        "spns_fft_destroy": (I32, [P]),
        "spns_fft_r2c": (I32, [P, P, P]),
        "spns_fft_c2r": (I32, [P, P, P]),
        "spns_fft_normalize": (I32, [P, P, I32, I32, I32]),
        "spns_fft_synchronize": (I32, [P]),
        "spns_fft_last_error": (ctypes.c_char_p, [P]),
        "spns_batch_create": (I32, [I32, I32, I32, ctypes.POINTER(P)]),
        "spns_batch_destroy": (I32, [P]),
        "spns_batch_attach": (I32, [P, I32, P]),
        "spns_batch_detach": (I32, [P, I32]),
        "spns_batch_step": (I32, [P, ctypes.POINTER(I32)]),
        "spns_batch_query": (I32, [P, ctypes.POINTER(I32)]),
        "spns_batch_synchronize": (I32, [P]),
        "spns_batch_last_error": (ctypes.c_char_p, [P]),
        "spns_batch_slot_scalars": (I32, [P, I32, ctypes.POINTER(Scalars)]),
        "spns_batch_fetch": (I32, [P, I32, D, D]),
        "spns_host_alloc": (I32, [I64, ctypes.POINTER(P)]),
        "spns_host_free": (I32, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class SpnsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("spns error %d: %s" % (code, msg))
        self.code = code


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(Nx, Ny=None, Nz=None, Mx=None, My=None, Mz=None, nu=1e-2, C_Smag=0.0, C_At=1.0, Lx=1, Ly=1, Lz=1,
                omega=(0.0, 0.0, 0.0), isNotCubic=False, isSelfAdaptive=False, isCCC=False, precision=PREC_FP64,
                delta=0.0, fixed_dt=0.0, force_eps=None, force_kmin=1, force_kmax=2, device=0, use_graph=True,
                zpass_variant=0, gradient_invariants=False, keep_step_start=False):
    """spns_config for the arguments of Solver."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    c = Config()
    c.Nx, c.Ny, c.Nz = Nx, Ny, Nz
    c.Mx, c.My, c.Mz = Mx or 0, My or 0, Mz or 0
    c.nu, c.C_Smag, c.C_At = nu, C_Smag, C_At
    c.Lx, c.Ly, c.Lz = Lx, Ly, Lz
    c.omega_x, c.omega_y, c.omega_z = omega
    c.isNotCubic, c.isSelfAdaptive, c.isComplexConjugateCorrected = int(isNotCubic), int(isSelfAdaptive), int(isCCC)
    c.precision = precision
    c.delta, c.fixed_dt = delta, fixed_dt
    c.force_on = int(force_eps is not None)
    c.force_eps = 0.0 if force_eps is None else force_eps
    c.force_kmin, c.force_kmax = force_kmin, force_kmax
    c.device = device
    c.use_graph = int(use_graph)
    c.zpass_variant = zpass_variant
    c.gradient_invariants = int(gradient_invariants)
    c.keep_step_start = int(keep_step_start)
    return c


def device_bytes(**kw):
    """Device memory a Solver with these arguments allocates."""
    return int(load_library().spns_device_bytes(ctypes.byref(make_config(**kw))))


def copy_bandwidth(nbytes=4 << 30, reps=10, device=0):
    """Measured streaming-copy bandwidth (GB/s of bytes read + written) of a
    hand-written copy kernel between two buffers of nbytes (spns_copy_bandwidth)."""
    L = load_library()
    g = ctypes.c_double()
    rc = L.spns_copy_bandwidth(device, nbytes, reps, ctypes.byref(g))
    if rc:
        raise SpnsError(rc, L.spns_last_error(None).decode())
    return g.value


def device_memory(device=0):
    """(free, total) bytes of a HIP device (spns_device_memory)."""
    L = load_library()
    fr, tot = ctypes.c_int64(), ctypes.c_int64()
    rc = L.spns_device_memory(device, ctypes.byref(fr), ctypes.byref(tot))
    if rc:
        raise SpnsError(rc, L.spns_last_error(None).decode())
    return int(fr.value), int(tot.value)


class Solver:
    """One HIT run on the GPU.  Keyword arguments follow HIT::HIT
    (src/HIT_LES.h:23-29) after MAIN's length-factor expansion, plus
    precision, delta, fixed_dt and constant-power forcing options."""

    def __init__(self, Nx, Ny=None, Nz=None, Mx=None, My=None, Mz=None, nu=1e-2, C_Smag=0.0, C_At=1.0,
                 Lx=1, Ly=1, Lz=1, omega=(0.0, 0.0, 0.0), isNotCubic=False, isSelfAdaptive=False,
                 isCCC=False, precision=PREC_FP64, delta=0.0, fixed_dt=0.0, force_eps=None,
                 force_kmin=1, force_kmax=2, device=0, use_graph=True, zpass_variant=0, gradient_invariants=False,
                 keep_step_start=False):
        self.L = load_library()
        self.args = dict(Nx=Nx, Ny=Ny, Nz=Nz, Mx=Mx, My=My, Mz=Mz, nu=nu, C_Smag=C_Smag, C_At=C_At, Lx=Lx, Ly=Ly, Lz=Lz,
                         omega=omega, isNotCubic=isNotCubic, isSelfAdaptive=isSelfAdaptive, isCCC=isCCC,
                         precision=precision, delta=delta, fixed_dt=fixed_dt, force_eps=force_eps, force_kmin=force_kmin,
                         force_kmax=force_kmax, device=device, use_graph=use_graph, zpass_variant=zpass_variant,
                         gradient_invariants=gradient_invariants, keep_step_start=keep_step_start)
        c = make_config(Nx, Ny, Nz, Mx, My, Mz, nu, C_Smag, C_At, Lx, Ly, Lz, omega, isNotCubic, isSelfAdaptive,
                        isCCC, precision, delta, fixed_dt, force_eps, force_kmin, force_kmax, device, use_graph,
                        zpass_variant, gradient_invariants, keep_step_start)
        Ny = Nx if Ny is None else Ny
        Nz = Nx if Nz is None else Nz
        self.cfg = c
        h = ctypes.c_void_p()
        rc = self.L.spns_create(ctypes.byref(c), ctypes.byref(h))
        if rc:
            raise SpnsError(rc, self.L.spns_last_error(None).decode())
        self.h = h
        self.nmodes = self.L.spns_nmodes(h)
        self.nbins = self.L.spns_nbins(h)
        self.last_rad_max = self.nbins - 1
        self.last_rad = self.L.spns_last_rad(h)
        self.shape = ((3 * Nx) // 2 if not Mx else Mx, (3 * Ny) // 2 if not My else My, (3 * Nz) // 2 if not Mz else Mz)
        self.nk = (2 * (Ny // 2) + 1, 2 * (Nx // 2) + 1, Nz // 2 + 1)

    def _chk(self, rc):
        if rc:
            raise SpnsError(rc, self.L.spns_last_error(self.h).decode())

    def reconfigure(self, **kw):
        """Reuses this run's device buffers for another run of the same shape
        class (spns_reconfigure): new physics (nu, C_Smag, C_At, omega,
        fixed_dt, forcing, delta), state and history cleared.  The grid,
        precision and options must stay; set an initial field next."""
        args = dict(self.args)
        args.update(kw)
        c = make_config(**args)
        self._chk(self.L.spns_reconfigure(self.h, ctypes.byref(c)))
        self.args = args
        self.cfg = c

    def close(self):
        if getattr(self, "h", None):
            self.L.spns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- initial fields (HIT_init) ---
    def set_state(self, state):
        s = np.ascontiguousarray(state, dtype=np.complex128).reshape(3, self.nmodes)
        self._chk(self.L.spns_set_state(self.h, _dp(s.view(np.float64))))

    def get_state(self):
        out = np.zeros((3, self.nmodes), dtype=np.complex128)
        self._chk(self.L.spns_get_state(self.h, _dp(out.view(np.float64))))
        return out

    def get_state_planes(self, i2_first, n_i2):
        """State on the i2 planes [i2_first, i2_first + n_i2): [3][n_i2][nk1][nk3]."""
        out = np.zeros((3, n_i2, self.nk[1], self.nk[2]), dtype=np.complex128)
        self._chk(self.L.spns_get_state_planes(self.h, i2_first, n_i2, _dp(out.view(np.float64))))
        return out

    def get_rhist(self):
        out = np.zeros((3, self.nmodes), dtype=np.complex128)
        self._chk(self.L.spns_get_rhist(self.h, _dp(out.view(np.float64))))
        return out

    def init_from_physical(self, u, v, w):
        nf = u.shape
        u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
        ekf, ensf = ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.spns_init_from_physical(self.h, _dp(u), _dp(v), _dp(w), nf[0], nf[1], nf[2],
                                                 ctypes.byref(ekf), ctypes.byref(ensf)))
        return ekf.value, ensf.value

    def init_k41(self, nranks=4):
        self._chk(self.L.spns_init_k41(self.h, nranks))

    def init_random(self, seed=0, k0=4.0, urms=1.0):
        self._chk(self.L.spns_init_random(self.h, seed, k0, urms))

    def finish_init(self, isReLambda=False):
        self._chk(self.L.spns_finish_init(self.h, int(isReLambda)))

    # --- stepping / forcing ---
    def step(self, n=1):
        self._chk(self.L.spns_step(self.h, n))

    def synchronize(self):
        self._chk(self.L.spns_synchronize(self.h))

    def query(self):
        """True when the steps enqueued by the last step() call have finished.
        Raises SpnsError (code SPNS_ENONFINITE) once they have, if the run
        diverged (NaN/Inf velocity, energy, enstrophy or time step)."""
        d = ctypes.c_int32()
        self._chk(self.L.spns_query(self.h, ctypes.byref(d)))
        return bool(d.value)

    def last_step_ms(self):
        ms = ctypes.c_double()
        self._chk(self.L.spns_last_step_ms(self.h, ctypes.byref(ms)))
        return ms.value

    PASSES = ["clear", "inv_y", "inv_x", "zpass", "fwd_x", "fwd_y", "epilogue", "finalize"]

    def profile_passes(self, nsteps):
        """Mean device ms of each pass over nsteps eagerly launched steps."""
        ms = np.zeros(8)
        self._chk(self.L.spns_profile_passes(self.h, nsteps, _dp(ms)))
        return dict(zip(self.PASSES, ms.tolist()))

    def recalc_timestep(self):
        self._chk(self.L.spns_recalculate_timestep(self.h))

    def check_energy(self, atol=1e-13):
        """HIT::Check_PhysFour_Energy: (EkReal, EkFour, EkSpNS) of the current
        state; raises SpnsError (SPNS_ECHECK) if they differ by more than atol."""
        er, ef, es = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        rc = self.L.spns_check_phys_four_energy(self.h, atol, ctypes.byref(er), ctypes.byref(ef), ctypes.byref(es))
        self._chk(rc)
        return er.value, ef.value, es.value

    def antitransform(self):
        """Refresh the CFL maxima from the current state (no copy-out)."""
        self._chk(self.L.spns_recalculate_velocity_antitransform(self.h))

    def set_nu(self, nu):
        self._chk(self.L.spns_set_nu(self.h, nu))

    def forcing_relambda(self, relambda):
        self._chk(self.L.spns_forcing_reynolds_lambda(self.h, relambda))

    def relambda(self):
        r = ctypes.c_double()
        self._chk(self.L.spns_recalculate_reynolds_lambda(self.h, ctypes.byref(r)))
        return r.value

    def forcing_energy_cascade(self, ek, last_input_rad, nullify=False):
        ek = np.ascontiguousarray(ek, dtype=np.float64)
        self._chk(self.L.spns_forcing_energy_cascade(self.h, _dp(ek), last_input_rad, int(nullify)))

    def ccc(self):
        self._chk(self.L.spns_complex_conjugate_correction(self.h))

    # --- diagnostics ---
    def invariant_spectra(self):
        """Shell spectra of QG, QS, RG, RS, V2 of the velocity gradients at the
        start of the last step: array [5][nbins] (needs gradient_invariants)."""
        out = np.zeros((5, self.nbins))
        self._chk(self.L.spns_invariant_spectra(self.h, self.nbins, _dp(out)))
        return out

    def spectrum(self):
        ek = np.zeros(self.nbins)
        tot = ctypes.c_double()
        self._chk(self.L.spns_energy_cascade(self.h, _dp(ek), self.nbins, ctypes.byref(tot)))
        return ek

    def energy(self):
        return self.scalars()["ek"]

    def enstrophy(self):
        return self.scalars()["om"]

    def pass_kernels(self):
        """(z-pass kernel: 2 staged, 1 block per pencil, 0 wave per pencil;
        fused epilogue: 0 separate kernels, 1 the forward y pass fused with the
        epilogue, 2 that kernel also running the next step's inverse y pass)."""
        z, f = ctypes.c_int32(), ctypes.c_int32()
        self._chk(self.L.spns_pass_kernels(self.h, ctypes.byref(z), ctypes.byref(f)))
        return z.value, f.value

    def scalars(self):
        s = Scalars()
        self._chk(self.L.spns_get_scalars(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Scalars._fields_}

    def velocity_physical(self):
        Mx, My, Mz = self.shape
        out = np.zeros((3, Mx, My, Mz))
        self._chk(self.L.spns_velocity_physical(self.h, _dp(out)))
        return out

    def velocity_physical_step_start(self):
        """Physical field of the state at the start of the last step (what the
        reference's u,v,w arrays hold between steps); CFL maxima untouched."""
        Mx, My, Mz = self.shape
        out = np.zeros((3, Mx, My, Mz))
        self._chk(self.L.spns_velocity_physical_step_start(self.h, _dp(out)))
        return out


class Fft:
    """FFT_package drop-in on device buffers (src/FFTW_package.h:21-71).
    Buffers are device pointers (e.g. torch tensors' data_ptr())."""

    def __init__(self, Mx, My, Mz, precision=PREC_FP64, device=0):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.spns_fft_create(Mx, My, Mz, precision, device, ctypes.byref(h))
        if rc:
            raise SpnsError(rc, self.L.spns_fft_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc:
            raise SpnsError(rc, self.L.spns_fft_last_error(self.h).decode())

    def r2c(self, rin_ptr, cout_ptr):
        self._chk(self.L.spns_fft_r2c(self.h, ctypes.c_void_p(rin_ptr), ctypes.c_void_p(cout_ptr)))

    def c2r(self, cin_ptr, rout_ptr):
        self._chk(self.L.spns_fft_c2r(self.h, ctypes.c_void_p(cin_ptr), ctypes.c_void_p(rout_ptr)))

    def normalize(self, c_ptr, Nx=0, Ny=0, Nz=0):
        self._chk(self.L.spns_fft_normalize(self.h, ctypes.c_void_p(c_ptr), Nx, Ny, Nz))

    def synchronize(self):
        self._chk(self.L.spns_fft_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.spns_fft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched stepping of up to nslots runs of one shape class
    (spns_batch_*): each pass of a step is one kernel launch over all
    attached runs; a chunk of up to `chunk` steps is one captured graph."""

    def __init__(self, nslots, chunk, device=0):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.spns_batch_create(device, nslots, chunk, ctypes.byref(h))
        if rc:
            raise SpnsError(rc, self.L.spns_batch_last_error(None).decode())
        self.h = h
        self.nslots, self.chunk = nslots, chunk
        self.solvers = [None] * nslots

    def _chk(self, rc):
        if rc:
            raise SpnsError(rc, self.L.spns_batch_last_error(self.h).decode())

    def attach(self, slot, solver):
        self._chk(self.L.spns_batch_attach(self.h, slot, solver.h))
        self.solvers[slot] = solver

    def detach(self, slot):
        self._chk(self.L.spns_batch_detach(self.h, slot))
        self.solvers[slot] = None

    def free_slots(self):
        return [i for i, s in enumerate(self.solvers) if s is None]

    def step(self, nsteps):
        """nsteps: per-slot step counts (0..chunk)."""
        n = (ctypes.c_int32 * self.nslots)(*[int(v) for v in nsteps])
        self._chk(self.L.spns_batch_step(self.h, n))

    def query(self):
        d = ctypes.c_int32()
        self._chk(self.L.spns_batch_query(self.h, ctypes.byref(d)))
        return bool(d.value)

    def slot_scalars(self, slot):
        """Scalars of a slot after the last chunk (no device read; the batch
        must be idle).  Raises SpnsError (SPNS_ENONFINITE) if it diverged."""
        sc = Scalars()
        self._chk(self.L.spns_batch_slot_scalars(self.h, slot, ctypes.byref(sc)))
        return {k: getattr(sc, k) for k, _ in Scalars._fields_}

    def fetch(self, slot, state=None, ek=None):
        """Queues copies of a slot's state and spectrum into (pinned) host
        arrays; complete once query() is True."""
        sp = _dp(state.view(np.float64).reshape(-1)) if state is not None else None
        ep = _dp(ek) if ek is not None else None
        self._chk(self.L.spns_batch_fetch(self.h, slot, sp, ep))

    def synchronize(self):
        self._chk(self.L.spns_batch_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.spns_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pinned:
    """A numpy array in pinned host memory (spns_host_alloc), for
    asynchronous device-to-host copies.  close() frees it; copy the data out
    first."""

    def __init__(self, shape, dtype=np.float64):
        self.L = load_library()
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        rc = self.L.spns_host_alloc(nbytes, ctypes.byref(p))
        if rc:
            raise SpnsError(rc, self.L.spns_last_error(None).decode())
        self.ptr = p
        buf = (ctypes.c_char * nbytes).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            self.L.spns_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
