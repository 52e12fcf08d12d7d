"""Python interface of this project's network libraries: KROME-X's
`Network.solve` (runtime/interfaces/python/krome.py:136-542) over the C
ABI of libkrome_<network>.so (include/krome.h), whose batched solves run on
the GPU (hand-written HIP kernels for gfx950).

    import krome
    net = krome.load("primordial3")
    opts = net.options(evolve_T=True, rtol=1e-8)
    res = net.solve(n, Tgas=100.0, dt=3.15e13, par={"tff": 1e12}, opts=opts)

Array layouts follow the C interface in numpy's row-major order: n[ncell,
nspec], par[ncell, npar]; a single cell may be 1-D arrays and scalars and
comes back the same way.  solve() returns NEW arrays, the inputs are not
modified, and raises KromeError when a cell fails (check=False: inspect
result.flags).  solve_device() works on torch tensors resident on the GPU
(species-major n[nspec, ncell], the coalesced layout).

There is no CPU path: loading a library that was not built, or a machine
without a HIP device, fails loudly.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

__all__ = ["load", "load_multi", "Network", "Multi", "Options", "Result",
           "KromeError"]

HERE = Path(__file__).resolve().parent
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_L = ctypes.POINTER(ctypes.c_long)
_ENUMS = {"jac": ("auto", "analytic", "fd"),
          "linsol": ("auto", "dense", "lapack", "sparse"),
          "path": ("auto", "percell", "device"),
          "device_kernel": ("plain", "persistent", "lanes")}
_KEYS = ("evolve_T", "jac", "linsol", "rtol", "atol_n", "atol_T", "atol_u",
         "max_steps", "max_err_fails", "max_conv_fails", "max_restarts",
         "clip_negative", "error_report", "renormalize_x", "conserve",
         "path", "device_cache", "device_memory", "min_device",
         "min_interleave", "device_kernel", "mixed_precision",
         "queue_chunk", "waves_per_cu", "lane_block",
         "cells_per_workgroup", "checkpoint_steps", "checkpoint_memory")
LAYOUT_SPECIES_MAJOR = 0
LAYOUT_CELL_MAJOR = 1


class KromeError(RuntimeError):
    pass


def _dptr(a):
    """double* of a float64 numpy array (None: NULL)"""
    if a is None:
        return None
    return ctypes.cast(a.ctypes.data, _D)


class _Batch:
    """one batch call's arrays in the C interface's conventions.

    The cell rows (n or x, [ncell, nspec] row-major, or one cell as
    [nspec]) are copied into a fresh C-ordered float64 block: the C
    function updates that block in place, so the caller's array never
    changes.  Per-cell scalars (Tgas, dt, rho) are expanded to one value
    per cell.  take() gives an output back in the shape of the input: a
    single cell given 1-D comes back 1-D, its scalars as scalars."""

    def __init__(self, net, rows, what="n"):
        a = np.asarray(rows, dtype=np.float64)
        if a.ndim not in (1, 2) or a.shape[-1] != net.nspec:
            raise KromeError(
                "%s: expected %d values per cell ([%d] or [ncell, %d]), got "
                "shape %s" % (what, net.nspec, net.nspec, net.nspec,
                              a.shape))
        self.one = a.ndim == 1
        self.rows = np.array(a.reshape(-1, net.nspec), order="C")
        self.ncell = self.rows.shape[0]

    def column(self, v, what):
        """one float64 per cell from a scalar or an [ncell] array"""
        out = np.empty(self.ncell)
        try:
            out[:] = np.asarray(v, dtype=np.float64)
        except ValueError:
            raise KromeError("%s: expected a scalar or %d values, got shape "
                             "%s" % (what, self.ncell, np.shape(v))) from None
        return out

    def take(self, a):
        if a is None or not self.one:
            return a
        return a[0]


class _Extra(ctypes.Structure):
    """krome_cells: the optional per-cell arrays of a multi-cell call"""
    _fields_ = [("Phi", ctypes.c_void_p), ("Td", ctypes.c_void_p),
                ("adust", ctypes.c_void_p)]


def _extra(F):
    """a krome_cells pointer holding the fields F (None: no extra)"""
    if F is None:
        return None
    return ctypes.byref(_Extra(F.ctypes.data, None, None))


def lib_path(network):
    return HERE / ("libkrome_%s.so" % network)


def load(network="primordial3", path=None):
    """the Network of a built library (by network name, or its path)"""
    p = Path(path) if path else lib_path(network)
    if not p.exists():
        raise KromeError("%s is not built (python3 krome-x_amd/build.py)" % p)
    # one HIP runtime per process: when torch is importable its runtime
    # (soname libamdhip64.so.7) is loaded first and serves this library
    # too, so torch tensors and this library's kernels share the device
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return Network(ctypes.CDLL(str(p)))


class _MultiBatch(ctypes.Structure):
    """krome_multi_batch (include/krome_multi.h)"""
    _fields_ = [("network", ctypes.c_int), ("ncell", ctypes.c_int),
                ("n", ctypes.c_void_p), ("Tgas", ctypes.c_void_p),
                ("dt", ctypes.c_void_p), ("par", ctypes.c_void_p),
                ("flags", ctypes.c_void_p), ("evolve_T", ctypes.c_int),
                ("rtol", ctypes.c_double), ("atol_n", ctypes.c_double),
                ("atol_T", ctypes.c_double), ("max_steps", ctypes.c_long),
                ("jac", ctypes.c_int), ("stats", ctypes.c_void_p)]


def load_multi(name="multi", path=None):
    """a library of several networks in one persistent launch
    (libkrome_<name>.so, include/krome_multi.h): "multi" (primordial,
    primordial3) or "multi_c3" (primordial, primordialZ, cloud)"""
    p = Path(path) if path else lib_path(name)
    if not p.exists():
        raise KromeError("%s is not built (python3 krome-x_amd/build.py)" % p)
    try:
        import torch  # noqa: F401  (the same HIP runtime as torch)
    except ImportError:
        pass
    return Multi(ctypes.CDLL(str(p)))


class Multi:
    """several networks' cells in one persistent launch (this build's
    extension, no counterpart in the reference): networks[i] names
    network i of the library"""

    def __init__(self, lib):
        self._lib = lib
        lib.krome_multi_nnet.restype = ctypes.c_int
        lib.krome_multi_network_name.restype = ctypes.c_char_p
        lib.krome_multi_network_name.argtypes = [ctypes.c_int]
        for f in ("evolve_T", "lockstep", "mode", "nspec", "npar"):
            getattr(lib, "krome_multi_" + f).restype = ctypes.c_int
            getattr(lib, "krome_multi_" + f).argtypes = [ctypes.c_int]
        lib.krome_multi_set_launch.restype = ctypes.c_int
        lib.krome_multi_set_launch.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.krome_multi_solve_device.restype = ctypes.c_int
        lib.krome_multi_solve_device.argtypes = [
            ctypes.POINTER(_MultiBatch), ctypes.c_int, ctypes.c_int, _D]
        count = lib.krome_multi_nnet()
        self.networks = [lib.krome_multi_network_name(i).decode()
                         for i in range(count)]
        self.evolve_T = [lib.krome_multi_evolve_T(i) for i in range(count)]
        self.lockstep = [lib.krome_multi_lockstep(i) for i in range(count)]
        # 1 the lockstep kernel, 2 a wave per cell, 0 the lane queue
        self.mode = [lib.krome_multi_mode(i) for i in range(count)]

    def set_launch(self, waves_per_cu=0, queue_chunk=0):
        """the launch shape of the following solves (0: the defaults)"""
        if self._lib.krome_multi_set_launch(int(waves_per_cu),
                                            int(queue_chunk)) != 0:
            raise KromeError("launch shape out of range")

    def solve_device(self, batches, mixed_precision=1):
        """batches: dicts of network (name), n [nspec, ncell], Tgas, dt
        [ncell], par [ncell, npar], flags (int32 [ncell]), all contiguous
        float64 torch tensors on the GPU (updated in place), and evolve_T,
        rtol, atol_n, atol_T, max_steps, jac, and optionally stats (int64
        [ncell, 9] on the GPU: the cells' counters).  Returns (the C
        function's flag, the launch's kernel ms)"""
        arr = (_MultiBatch * len(batches))()
        for k, b in enumerate(batches):
            for key in ("n", "Tgas", "dt", "par", "flags"):
                t = b[key]
                if not t.is_cuda or not t.is_contiguous():
                    raise KromeError("solve_device needs contiguous GPU "
                                     "tensors (%s)" % key)
            arr[k] = _MultiBatch(
                self.networks.index(b["network"]), int(b["Tgas"].shape[0]),
                b["n"].data_ptr(), b["Tgas"].data_ptr(), b["dt"].data_ptr(),
                b["par"].data_ptr(), b["flags"].data_ptr(),
                int(b.get("evolve_T", 0)), float(b.get("rtol", 1e-8)),
                float(b.get("atol_n", 1e-20)), float(b.get("atol_T", 1e-6)),
                int(b.get("max_steps", 100000)), int(b.get("jac", 0)),
                b["stats"].data_ptr() if b.get("stats") is not None
                else None)
        ms = ctypes.c_double()
        r = self._lib.krome_multi_solve_device(
            arr, len(batches), int(mixed_precision),
            ctypes.cast(ctypes.byref(ms), _D))
        return r, ms.value


class Network:
    """one network's library: its species, parameters and solves"""

    def __init__(self, lib):
        self._lib = lib
        L = lib
        for name, res, args in (
                ("krome_nspec", ctypes.c_int, []),
                ("krome_nrea", ctypes.c_int, []),
                ("krome_npar", ctypes.c_int, []),
                ("krome_network_name", ctypes.c_char_p, []),
                ("krome_version", ctypes.c_char_p, []),
                ("krome_species_name", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_param_name", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_param_required_by", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_flag_message", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_flag_name", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_params_default", None, [_D]),
                ("krome_opts_new", ctypes.c_void_p, []),
                ("krome_opts_free", None, [ctypes.c_void_p]),
                ("krome_opts_set", ctypes.c_int,
                 [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]),
                ("krome_opts_get", ctypes.c_int,
                 [ctypes.c_void_p, ctypes.c_char_p, _D]),
                ("krome_opts_set_atol_species", ctypes.c_int,
                 [ctypes.c_void_p, _D]),
                ("krome_opts_set_rtol_species", ctypes.c_int,
                 [ctypes.c_void_p, _D]),
                ("krome_opts_get_atol_species", ctypes.c_int,
                 [ctypes.c_void_p, _D]),
                ("krome_opts_get_rtol_species", ctypes.c_int,
                 [ctypes.c_void_p, _D]),
                ("krome_has_device", ctypes.c_int, []),
                ("krome_device_initialize", ctypes.c_int, []),
                ("krome_device_finalize", ctypes.c_int, []),
                ("krome_solve_cells", ctypes.c_int,
                 [ctypes.c_int, _D, _D, _D, _D, ctypes.c_int,
                  ctypes.c_void_p, _I, ctypes.c_void_p]),
                ("krome_solve_cells_x", ctypes.c_int,
                 [ctypes.c_int, _D, _D, _D, _D, _D, ctypes.c_int,
                  ctypes.c_void_p, _I, ctypes.c_void_p]),
                ("krome_solve_cells_device", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_long,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p]),
                ("krome_solve_cells_x_device", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_long, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]),
                ("krome_solve_cells_device_stats", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_long,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]),
                ("krome_device_last_kernel_ms", ctypes.c_int, [_D]),
                ("krome_solve_cells_device_noalloc", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_long,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]),
                ("krome_adjoint_cells_device", ctypes.c_int,
                 [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int] +
                 [ctypes.c_void_p] * 9),
                ("krome_get_rates", ctypes.c_int,
                 [ctypes.c_int, _D, _D, _D, ctypes.c_int, _D,
                  ctypes.c_void_p, _D]),
                ("krome_get_fluxes", ctypes.c_int,
                 [ctypes.c_int, _D, _D, _D, ctypes.c_int, _D,
                  ctypes.c_void_p, _D]),
                ("krome_nheat", ctypes.c_int, []),
                ("krome_ncool", ctypes.c_int, []),
                ("krome_heating_name", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_cooling_name", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_get_thermal", ctypes.c_int,
                 [ctypes.c_int, _D, _D, _D, ctypes.c_int, _D,
                  ctypes.c_void_p, _D, _D, _D, _D, _D]),
                ("krome_species_mass", ctypes.c_double, [ctypes.c_int]),
                ("krome_species_charge", ctypes.c_int, [ctypes.c_int]),
                ("krome_nelem", ctypes.c_int, []),
                ("krome_element_name", ctypes.c_char_p, [ctypes.c_int]),
                ("krome_species_natoms", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_int]),
                ("krome_get_element_totals", ctypes.c_int,
                 [ctypes.c_int, _D, _D]),
                ("krome_get_element_totals_x", ctypes.c_int,
                 [ctypes.c_int, _D, _D]),
                ("krome_conserve_cells", ctypes.c_int,
                 [ctypes.c_int, _D, _D, ctypes.c_int, _D, ctypes.c_int, _I]),
                ("krome_conserve_cells_x", ctypes.c_int,
                 [ctypes.c_int, _D, _D, ctypes.c_int, _D, ctypes.c_int, _I]),
                ("krome_conserve_cells_device", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_long,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_void_p]),
                ("krome_conserve_cells_x_device", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_long,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_void_p]),
                ("krome_opts_set_photo", ctypes.c_int,
                 [ctypes.c_void_p, ctypes.c_int, _D, _D]),
                ("krome_opts_nbins", ctypes.c_int, [ctypes.c_void_p]),
                ("krome_opts_set_photo_alpha", ctypes.c_int,
                 [ctypes.c_void_p, ctypes.c_double]),
                ("krome_opts_set_photo_broadening", ctypes.c_int,
                 [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
                ("krome_opts_set_photo_dust_opacity", ctypes.c_int,
                 [ctypes.c_void_p, _D]),
                ("krome_opts_photo_sigma", ctypes.c_int,
                 [ctypes.c_void_p, _D]),
                ("krome_nphoto", ctypes.c_int, []),
                ("krome_photo_reaction", ctypes.c_int, [ctypes.c_int]),
                ("krome_photo_reactant", ctypes.c_int, [ctypes.c_int]),
                ("krome_photo_pdiss", ctypes.c_double, [ctypes.c_int]),
                ("krome_photo_xsec", ctypes.c_double,
                 [ctypes.c_int, ctypes.c_double]),
                ("krome_bins_lin", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D]),
                ("krome_bins_log", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D,
                  _D])):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.nspec = L.krome_nspec()
        self.nrea = L.krome_nrea()
        self.npar = L.krome_npar()
        self.name = L.krome_network_name().decode()
        self.species = [L.krome_species_name(i).decode()
                        for i in range(self.nspec)]
        self.params = [L.krome_param_name(i).decode()
                       for i in range(self.npar)]
        self.required = {p: L.krome_param_required_by(i).decode()
                         for i, p in enumerate(self.params)
                         if L.krome_param_required_by(i)}
        self._pidx = {p: i for i, p in enumerate(self.params)}
        self._opts = None
        # species and element metadata (the reference binding's
        # krome.py:334-345)
        self.masses = np.array([L.krome_species_mass(i)
                                for i in range(self.nspec)])  # [g]
        self.charges = np.array([L.krome_species_charge(i)
                                 for i in range(self.nspec)])
        self.elements = tuple(L.krome_element_name(e).decode()
                              for e in range(L.krome_nelem()))
        self.heating_terms = tuple(L.krome_heating_name(i).decode()
                                   for i in range(L.krome_nheat()))
        self.cooling_terms = tuple(L.krome_cooling_name(i).decode()
                                   for i in range(L.krome_ncool()))
        self.composition = np.array(
            [[L.krome_species_natoms(i, e)
              for e in range(len(self.elements))]
             for i in range(self.nspec)], dtype=int).reshape(
                 self.nspec, len(self.elements))

    def idx(self, name):
        return self.species.index(name)

    def params_default(self):
        p = np.zeros(self.npar)
        self._lib.krome_params_default(_dptr(p))
        return p

    def param_rows(self, ncell, **values):
        rows = np.tile(self.params_default(), (ncell, 1))
        for k, v in values.items():
            rows[:, self._pidx[k]] = np.broadcast_to(np.asarray(v, float),
                                                     (ncell,))
        return rows

    def options(self, **values):
        return Options(self, **values)

    def flag_message(self, flag):
        return self._lib.krome_flag_message(int(flag)).decode()

    def flag_name(self, flag):
        return self._lib.krome_flag_name(int(flag)).decode()

    # -- solves ----------------------------------------------------------

    def solve(self, n, Tgas, dt, par=None, Phi=None, opts=None, check=True,
              Td=None, adust=None):
        """integrate every cell over its own dt [s] (krome_solve_cells);
        Phi: the photon flux per bin [cm^-2 s^-1], needed once
        opts.set_photo() was called: Phi[nbins] for every cell or
        Phi[ncell, nbins].  Td and adust (the reference binding's dust-bin
        arguments) must be None: this build has no dust bins"""
        return self._solve("krome_solve_cells", n, (), Tgas, dt, par, Phi,
                           opts, check, Td, adust)

    def solve_x(self, x, rho, Tgas, dt, par=None, Phi=None, opts=None,
                check=True, Td=None, adust=None):
        """solve() with mass fractions x and the gas density rho [g cm^-3]
        (krome_solve_cells_x)"""
        return self._solve("krome_solve_cells_x", x, (("rho", rho),), Tgas,
                           dt, par, Phi, opts, check, Td, adust)

    def _solve(self, fname, rows, lead, Tgas, dt, par, Phi, opts, check, Td,
               adust):
        """the host-array solves: fname(ncell, rows, *lead, Tgas, dt, par,
        stride, extra, flags, opts) on a private copy of the inputs"""
        if Td is not None or adust is not None:
            raise KromeError("Td / adust need dust bins, which this build "
                             "does not have")
        b = _Batch(self, rows)
        opts = self._opts_or_default(opts)
        cols = [b.column(v, what) for what, v in
                (*lead, ("Tgas", Tgas), ("dt", dt))]
        P, stride = self._param_block(par, b.ncell)
        F = self._phi(Phi, b.ncell, opts)
        flags = np.zeros(b.ncell, dtype=np.intc)
        ierr = getattr(self._lib, fname)(
            b.ncell, _dptr(b.rows), *(_dptr(c) for c in cols), _dptr(P),
            stride, _extra(F), flags.ctypes.data_as(_I), opts._h)
        if check and ierr != 0:
            self._raise_failed(fname, ierr, flags)
        return Result(b.take(b.rows), b.take(cols[-2]), b.take(F),
                      b.take(flags), ierr)

    def solve_device(self, n, Tgas, dt, par, opts=None, flags=None,
                     stats=None, sync=True, Phi=None):
        """device-resident solve on torch tensors: n[nspec, ncell]
        (species-major, float64), Tgas, dt [ncell], par[ncell, npar], all
        on the GPU; updated in place.  flags (int32 [ncell]) and stats
        (int64 [ncell, 9]) optional device tensors; Phi[ncell, nbins] the
        photon fields on the GPU (krome_solve_cells_device with
        extra->Phi; no stats then).  Returns the C function's flag."""
        ncell = int(Tgas.shape[0])
        for t in (n, Tgas, dt, par):
            if not t.is_cuda or not t.is_contiguous():
                raise KromeError("solve_device needs contiguous GPU tensors")
        if tuple(n.shape) != (self.nspec, ncell):
            raise KromeError("n must be [nspec, ncell]")
        if tuple(par.shape) != (ncell, self.npar):
            raise KromeError("par must be [ncell, npar]")
        opts = self._opts_or_default(opts)
        if Phi is not None:
            if stats is not None:
                raise KromeError("no per-cell counters with a field")
            if (not Phi.is_cuda or not Phi.is_contiguous()
                    or tuple(Phi.shape) != (ncell, opts.nbins)):
                raise KromeError("Phi must be a contiguous GPU tensor "
                                 "[ncell, nbins]")
            ext = _Extra(Phi.data_ptr(), None, None)
            r = self._lib.krome_solve_cells_device(
                ncell, n.data_ptr(), LAYOUT_SPECIES_MAJOR, ncell,
                Tgas.data_ptr(), dt.data_ptr(), par.data_ptr(), self.npar,
                ctypes.byref(ext),
                flags.data_ptr() if flags is not None else None, None,
                opts._h)
            if sync:
                import torch
                torch.cuda.synchronize()
            return r
        if flags is None or stats is not None:
            import torch
            if flags is None:
                flags = torch.zeros(ncell, dtype=torch.int32,
                                    device=Tgas.device)
        r = self._lib.krome_solve_cells_device_stats(
            ncell, n.data_ptr(), LAYOUT_SPECIES_MAJOR, ncell, Tgas.data_ptr(),
            dt.data_ptr(), par.data_ptr(), self.npar, flags.data_ptr(),
            stats.data_ptr() if stats is not None else None, opts._h)
        if sync:
            import torch
            torch.cuda.synchronize()
        return r

    def solve_x_device(self, x, rho, Tgas, dt, par, opts=None, flags=None,
                       layout=None):
        """device-resident solve with mass fractions on torch tensors
        (krome_solve_cells_x_device): x [nspec, ncell] (species-major) or
        [ncell, nspec] (layout LAYOUT_CELL_MAJOR), rho, Tgas, dt [ncell],
        par [ncell, npar], float64 on the GPU; updated in place, failed
        cells keep x.  flags: optional int32 [ncell] device tensor.
        Returns the C function's flag"""
        if layout is None:
            layout = LAYOUT_SPECIES_MAJOR
        ncell = int(Tgas.shape[0])
        for t in (x, rho, Tgas, dt, par):
            if not t.is_cuda or not t.is_contiguous():
                raise KromeError("solve_x_device needs contiguous GPU "
                                 "tensors")
        shape = ((self.nspec, ncell) if layout == LAYOUT_SPECIES_MAJOR
                 else (ncell, self.nspec))
        if tuple(x.shape) != shape:
            raise KromeError("x must be %s" % (shape,))
        if tuple(par.shape) != (ncell, self.npar):
            raise KromeError("par must be [ncell, npar]")
        opts = self._opts_or_default(opts)
        return self._lib.krome_solve_cells_x_device(
            ncell, x.data_ptr(), rho.data_ptr(), layout, 0, Tgas.data_ptr(),
            dt.data_ptr(), par.data_ptr(), self.npar,
            None, flags.data_ptr() if flags is not None else None, None,
            opts._h)

    def adjoint_device(self, n, Tgas, dt, par, opts=None, lnk=None,
                       flags=None, gn=None, gT=None):
        """the discrete adjoint's solve on torch tensors (this build's
        extension, krome_adjoint_cells_device): n [nspec, ncell] species-
        major, Tgas, dt [ncell], par [ncell, npar], float64 on the GPU, n
        and Tgas updated in place; lnk [ncell, nrea] the log-rate offsets
        (rates k_r exp(lnk_r); None: 0).  With the cotangents gn [nspec,
        ncell] and/or gT [ncell] of the final state, also returns the
        gradients (g_lnk [ncell, nrea], g_n0 [nspec, ncell], g_T0 [ncell]
        or None at fixed Tgas).  Returns (flag, gradients or None, the
        peak device bytes of the adjoint)"""
        import torch
        ncell = int(Tgas.shape[0])
        ts = [n, Tgas, dt, par] + [t for t in (lnk, gn, gT) if t is not None]
        for t in ts:
            if (not t.is_cuda or not t.is_contiguous()
                    or t.dtype != torch.float64):
                raise KromeError("adjoint_device needs contiguous float64 "
                                 "GPU tensors")
        if tuple(n.shape) != (self.nspec, ncell):
            raise KromeError("n must be [nspec, ncell]")
        if tuple(par.shape) != (ncell, self.npar):
            raise KromeError("par must be [ncell, npar]")
        if lnk is not None and tuple(lnk.shape) != (ncell, self.nrea):
            raise KromeError("lnk must be [ncell, nrea]")
        opts = self._opts_or_default(opts)
        if flags is None:
            flags = torch.zeros(ncell, dtype=torch.int32, device=n.device)
        grads = None
        ptrs = [None] * 5
        if gn is not None or gT is not None:
            et = bool(opts["evolve_T"])
            g_lnk = torch.zeros((ncell, self.nrea), dtype=torch.float64,
                                device=n.device)
            g_n0 = torch.zeros_like(n)
            g_T0 = torch.zeros_like(Tgas) if et else None
            grads = (g_lnk, g_n0, g_T0)
            ptrs = [None if gn is None else gn.data_ptr(),
                    None if gT is None else gT.data_ptr(),
                    g_lnk.data_ptr(), g_n0.data_ptr(),
                    None if g_T0 is None else g_T0.data_ptr()]
        mem = ctypes.c_double(0.0)
        r = self._lib.krome_adjoint_cells_device(
            ncell, n.data_ptr(), Tgas.data_ptr(), dt.data_ptr(),
            par.data_ptr(), self.npar,
            None if lnk is None else lnk.data_ptr(), flags.data_ptr(),
            *ptrs, ctypes.addressof(mem), opts._h)
        return r, grads, mem.value

    # -- queries (the reference binding's rates / fluxes / thermal) --------

    def rates(self, n, Tgas, par=None, Phi=None, opts=None):
        """rate coefficients [..., nrea] at the states (krome_get_rates),
        the photo rates of the fields Phi over opts' photon bins"""
        return self._query(n, Tgas, par, Phi, opts, "krome_get_rates")

    def fluxes(self, n, Tgas, par=None, Phi=None, opts=None):
        """reaction fluxes k prod(n) [cm^-3 s^-1], [..., nrea]"""
        return self._query(n, Tgas, par, Phi, opts, "krome_get_fluxes")

    def _query_inputs(self, n, Tgas, par, Phi, opts):
        """(batch, Tgas column, parameter rows, stride, fields, opts
        handle) of a query; no opts: no photon bins"""
        b = _Batch(self, n)
        P, stride = self._param_block(par, b.ncell)
        F = None if opts is None else self._phi(Phi, b.ncell, opts)
        return (b, b.column(Tgas, "Tgas"), P, stride, F,
                None if opts is None else opts._h)

    def _query(self, n, Tgas, par, Phi, opts, fname):
        b, T, P, stride, F, h = self._query_inputs(n, Tgas, par, Phi, opts)
        out = np.zeros((b.ncell, self.nrea))
        r = getattr(self._lib, fname)(b.ncell, _dptr(b.rows), _dptr(T),
                                      _dptr(P), stride, _dptr(F), h,
                                      _dptr(out))
        if r != 0:
            raise KromeError("%s: %s (flag %d)" % (fname,
                                                   self.flag_message(r), r))
        return b.take(out)

    def thermal(self, n, Tgas, par=None, Phi=None, opts=None):
        """the thermal balance: dict of heat, cool [erg cm^-3 s^-1], the
        terms heats [..., nheat] and cools [..., ncool] (labels:
        heating_terms, cooling_terms) and dTdt [K/s]; the photoheating of
        the fields Phi over opts' photon bins"""
        b, T, P, stride, F, h = self._query_inputs(n, Tgas, par, Phi, opts)
        heat, cool, dTdt = (np.zeros(b.ncell) for _ in range(3))
        hs = np.zeros((b.ncell, max(len(self.heating_terms), 1)))
        cs = np.zeros((b.ncell, max(len(self.cooling_terms), 1)))
        r = self._lib.krome_get_thermal(b.ncell, _dptr(b.rows), _dptr(T),
                                        _dptr(P), stride, _dptr(F), h,
                                        _dptr(heat), _dptr(cool), _dptr(hs),
                                        _dptr(cs), _dptr(dTdt))
        if r != 0:
            raise KromeError("krome_get_thermal: %s (flag %d)"
                             % (self.flag_message(r), r))
        out = {"heat": heat, "cool": cool, "dTdt": dTdt,
               "heats": hs[:, :len(self.heating_terms)],
               "cools": cs[:, :len(self.cooling_terms)]}
        return {k: b.take(v) for k, v in out.items()}

    # -- element totals and consistent advection (the reference binding's
    #    krome.py:850-921) ------------------------------------------------

    def element_totals(self, n):
        """element totals [atoms cm^-3], [..., nelem] (labels: elements)"""
        return self._totals(n, "krome_get_element_totals")

    def element_totals_x(self, x):
        """element totals per gram of gas [atoms g^-1] from the mass
        fractions, [..., nelem] (legacy krome_conserveLinGetRef_x)"""
        return self._totals(x, "krome_get_element_totals_x")

    def _totals(self, v, fname):
        b = _Batch(self, v)
        nel = len(self.elements)
        out = np.zeros((b.ncell, max(nel, 1)))
        r = getattr(self._lib, fname)(b.ncell, _dptr(b.rows), _dptr(out))
        if r != 0:
            raise KromeError("%s: flag %d" % (fname, r))
        return b.take(out[:, :nel])

    def conserve(self, n, ref, par=None, flags=None, check=True):
        """consistent advection (legacy krome_conserveLin_x): the cells n
        ([nspec] or [ncell, nspec], cm^-3) moved back to the reference
        element totals ref ([nelem]: one row for all, or [ncell, nelem];
        element_totals of a reference cell) and to zero net charge
        (negatives clipped, the elements by the minimal relative
        correction, then the electrons from neutrality).  par: the
        parameters as for solve (only freeze directives read them).  flags:
        an int array [ncell] receiving each cell's flag (0,
        FLAG_CONSERVE_UNREACHED -112, the warning FLAG_CONSERVE_LARGE 2).
        Returns the corrected densities (the input is not modified); raises
        KromeError when a cell cannot reach its reference, unless
        check=False"""
        return self._conserve(n, ref, par, flags, check,
                              "krome_conserve_cells")

    def conserve_x(self, x, ref, par=None, flags=None, check=True):
        """conserve() for mass fractions, ref per gram (element_totals_x)"""
        return self._conserve(x, ref, par, flags, check,
                              "krome_conserve_cells_x")

    def _conserve(self, v, ref, par, flags, check, fname):
        b = _Batch(self, v)
        nel = len(self.elements)
        R = np.ascontiguousarray(ref, dtype=np.float64)
        shapes = {(nel,): 0, (b.ncell, nel): nel}
        if R.shape not in shapes:
            raise KromeError("ref: expected [%d] or [%d, %d], got shape %s"
                             % (nel, b.ncell, nel, R.shape))
        # the parameters only select freeze directives: REQUIRED ones may
        # stay unset here
        P, ps = self._param_block(par, b.ncell, require=False)
        fl = np.zeros(b.ncell, dtype=np.intc)
        r = getattr(self._lib, fname)(b.ncell, _dptr(b.rows), _dptr(R),
                                      shapes[R.shape], _dptr(P), ps,
                                      fl.ctypes.data_as(_I))
        if r == -1:
            raise KromeError("%s rejected its arguments" % fname)
        if flags is not None:
            flags[...] = fl.reshape(np.shape(flags))
        if check and r != 0:
            self._raise_failed(fname, r, np.where(fl < 0, fl, 0))
        return b.take(b.rows)

    def conserve_device(self, n, ref, flags=None, x=False, layout=None):
        """consistent advection on torch tensors on the GPU, in place: n
        [nspec, ncell] (species-major) or [ncell, nspec] (layout
        LAYOUT_CELL_MAJOR), float64; ref [nelem] or [ncell, nelem]; flags an
        int32 [ncell] device tensor (optional).  x: mass fractions.
        Returns the C function's flag"""
        if layout is None:
            layout = LAYOUT_SPECIES_MAJOR
        ncell = int(n.shape[1] if layout == LAYOUT_SPECIES_MAJOR
                    else n.shape[0])
        nel = len(self.elements)
        for t in (n, ref) + ((flags,) if flags is not None else ()):
            if not t.is_cuda or not t.is_contiguous():
                raise KromeError("conserve_device needs contiguous GPU "
                                 "tensors")
        rs = 0 if tuple(ref.shape) == (nel,) else nel
        if rs and tuple(ref.shape) != (ncell, nel):
            raise KromeError("ref must be [nelem] or [ncell, nelem]")
        f = (self._lib.krome_conserve_cells_x_device if x
             else self._lib.krome_conserve_cells_device)
        return f(ncell, n.data_ptr(), layout, 0, ref.data_ptr(), rs, None, 0,
                 flags.data_ptr() if flags is not None else None)

    # -- inputs (the krome_solve_cells conventions) -----------------------

    def _opts_or_default(self, opts):
        """opts, or the network's default handle (made on first use)"""
        if opts is None:
            if self._opts is None:
                self._opts = Options(self)
            opts = self._opts
        return opts

    def _phi(self, Phi, ncell, opts):
        """the fields as [ncell, nbins] rows (None when no bins are set)"""
        nb = opts.nbins
        if Phi is None:
            return None
        if nb == 0:
            raise KromeError("Phi given but no photon bins (opts.set_photo)")
        F = np.asarray(Phi, dtype=np.float64)
        if F.ndim == 1:
            F = np.tile(F, (ncell, 1))
        F = np.ascontiguousarray(F)
        if F.shape != (ncell, nb):
            raise KromeError("Phi must be [nbins] or [ncell, nbins]")
        return F

    # -- photon bins (krome.h photo functions) -----------------------------

    @property
    def nphoto(self):
        return self._lib.krome_nphoto()

    def photo_reactions(self):
        """[(reaction index, reactant index)] of the photo reactions"""
        return [(self._lib.krome_photo_reaction(r),
                 self._lib.krome_photo_reactant(r))
                for r in range(self.nphoto)]

    def photo_xsec(self, r, E):
        """cross section [cm^2] of photo reaction r at energies E [eV] (0
        outside its window)"""
        E = np.atleast_1d(np.asarray(E, dtype=np.float64))
        return np.array([self._lib.krome_photo_xsec(r, float(e))
                         for e in E])

    def bins_log(self, nbins, lower, upper):
        """(Elo, Ehi) of nbins logarithmic bins in [lower, upper] eV"""
        lo, hi = np.zeros(nbins), np.zeros(nbins)
        if self._lib.krome_bins_log(nbins, lower, upper, _dptr(lo),
                                    _dptr(hi)) != 0:
            raise KromeError("invalid binning")
        return lo, hi

    def bins_lin(self, nbins, lower, upper):
        """(Elo, Ehi) of nbins linear bins in [lower, upper] eV"""
        lo, hi = np.zeros(nbins), np.zeros(nbins)
        if self._lib.krome_bins_lin(nbins, lower, upper, _dptr(lo),
                                    _dptr(hi)) != 0:
            raise KromeError("invalid binning")
        return lo, hi

    def _param_block(self, par, ncell, require=True):
        """(rows, stride) for a C call: par None (every parameter at its
        default: rows None, stride 0), a dict {name: value or [ncell]
        values} over the defaults, one row [npar] shared by every cell
        (stride 0) or rows [ncell, npar] (stride npar).  With require, a
        REQUIRED parameter left NaN in any row is an error here, before
        the library sees it"""
        if par is None:
            P, stride = None, 0
        elif isinstance(par, dict):
            P, stride = self.param_rows(ncell, **par), self.npar
        else:
            P = np.ascontiguousarray(par, dtype=np.float64)
            stride = {(self.npar,): 0,
                      (ncell, self.npar): self.npar}.get(P.shape)
            if stride is None:
                raise KromeError("par: expected [%d] or [%d, %d], got shape "
                                 "%s" % (self.npar, ncell, self.npar,
                                         P.shape))
        if require and self.required:
            block = np.atleast_2d(self.params_default() if P is None else P)
            unset = [k for k in self.required
                     if np.isnan(block[:, self._pidx[k]]).any()]
            if unset:
                raise KromeError(
                    "REQUIRED parameter%s still NaN: %s" % (
                        "s" if len(unset) > 1 else "",
                        ", ".join("%s (read by %s)" % (k, self.required[k])
                                  for k in unset)))
        return P, stride

    def _raise_failed(self, what, ierr, flags):
        bad = np.flatnonzero(np.atleast_1d(flags))
        shown = bad[:16].tolist()
        more = " and %d more" % (bad.size - 16) if bad.size > 16 else ""
        raise KromeError("%s: %s (flag %d) in %d cell(s): %s%s; check=False "
                         "returns the result with its per-cell flags" % (
                             what, self.flag_message(ierr), ierr, bad.size,
                             shown, more))


class Options:
    """solver options (a krome_opts handle).  Keys: those of KROME-X's
    krome_opts_set (evolve_T, jac, linsol, rtol, atol_n, atol_T,
    max_steps, max_err_fails, max_conv_fails, max_restarts,
    clip_negative, error_report, renormalize_x, conserve, path,
    device_cache, device_memory, min_device, min_interleave) and this
    build's device_kernel ("plain" / "persistent"), mixed_precision,
    queue_chunk, waves_per_cu."""

    def __init__(self, net, **values):
        # the handle keeps a reference to its network: the library stays
        # loaded while the handle lives
        self._net, self._lib = net, net._lib
        self._h = self._lib.krome_opts_new()
        self.update(**values)

    def update(self, **values):
        """set several keys; returns self"""
        for key, value in values.items():
            self._set(key, value)
        return self

    def _set(self, key, value):
        if key not in _KEYS:
            raise KromeError("%r is not an option key (keys: %s)"
                             % (key, ", ".join(_KEYS)))
        names = _ENUMS.get(key)
        if names is not None and isinstance(value, str):
            if value not in names:
                raise KromeError("option %s takes %s, not %r"
                                 % (key, " / ".join(names), value))
            value = names.index(value)
        if self._lib.krome_opts_set(self._h, key.encode(), float(value)):
            raise KromeError("option %s: the library refused %r"
                             % (key, value))

    __setitem__ = _set

    def __getitem__(self, key):
        v = ctypes.c_double()
        if self._lib.krome_opts_get(self._h, key.encode(),
                                         ctypes.byref(v)) != 0:
            raise KromeError("unknown option %r" % key)
        return v.value

    def set_species_tolerances(self, atol=None, rtol=None):
        """per-species tolerances (krome_opts_set_atol_species /
        rtol_species): atol[nspec] [cm^-3] and rtol[nspec] in place of
        atol_n and rtol, a negative entry keeping the scalar; None leaves
        that kind as it is, "scalar" clears it"""
        ns = self._net.nspec
        for what, v in (("atol", atol), ("rtol", rtol)):
            if v is None:
                continue
            setter = getattr(self._lib, "krome_opts_set_%s_species" % what)
            if isinstance(v, str) and v == "scalar":
                setter(self._h, None)
                continue
            a = np.ascontiguousarray(v, dtype=np.float64)
            if a.shape != (ns,):
                raise KromeError("%s: expected %d values, got shape %s"
                                 % (what, ns, a.shape))
            if setter(self._h, _dptr(a)) != 0:
                raise KromeError("%s: NaN%s refused" % (
                    what, " or zero" if what == "atol" else ""))
        return self

    def species_tolerances(self):
        """(atol[nspec], rtol[nspec]): the tolerance each species gets"""
        a, r = np.zeros(self._net.nspec), np.zeros(self._net.nspec)
        self._lib.krome_opts_get_atol_species(self._h, _dptr(a))
        self._lib.krome_opts_get_rtol_species(self._h, _dptr(r))
        return a, r

    def set_photo(self, Elo, Ehi):
        """photon bins [eV edges]; Phi[..., nbins] is then required by
        solve() (the bins' cross sections are averaged once, here)"""
        lo = np.ascontiguousarray(Elo, dtype=np.float64)
        hi = np.ascontiguousarray(Ehi, dtype=np.float64)
        if lo.shape != hi.shape or lo.ndim != 1:
            raise KromeError("Elo and Ehi must be 1-D arrays of one shape")
        if self._lib.krome_opts_set_photo(self._h, lo.size, _dptr(lo),
                                               _dptr(hi)) != 0:
            raise KromeError("invalid photon bins (need 0 < Elo < Ehi)")
        return self

    def set_photo_alpha(self, alpha):
        """the in-bin photon spectrum dN/dE ~ E^-alpha (1: the same number
        per ln E); rebuilds the bins already set"""
        if self._lib.krome_opts_set_photo_alpha(self._h,
                                                     float(alpha)) != 0:
            raise KromeError("alpha must be finite")
        return self

    @property
    def nbins(self):
        return self._lib.krome_opts_nbins(self._h)

    def photo_sigma(self):
        """the bin-averaged cross sections [nphoto, nbins] [cm^2]"""
        s = np.zeros((max(self._net.nphoto, 1), self.nbins))
        if self._lib.krome_opts_photo_sigma(self._h, _dptr(s)) != 0:
            raise KromeError("no photon bins")
        return s[:self._net.nphoto]

    def close(self):
        if self._h is not None:
            self._lib.krome_opts_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Result:
    """what solve() returns, in the reference binding's field order: n,
    Tgas, Phi (the fields passed, held fixed during the solve; None without
    photon bins), flags (per cell, 0 = success), ierr (the flag of the
    first failing cell, 0 when all succeed), Td and adust (None: this build
    has no dust bins)"""

    def __init__(self, n, Tgas, Phi, flags, ierr, Td=None, adust=None):
        self.n = n
        self.Tgas = Tgas
        self.Phi = Phi
        self.flags = flags
        self.ierr = int(ierr)
        self.Td = Td
        self.adust = adust

    @property
    def ok(self):
        """True when every cell succeeded"""
        return not self.ierr

    def failed(self):
        """indices of the cells with a nonzero flag"""
        return np.flatnonzero(np.atleast_1d(self.flags) != 0)

    def __repr__(self):
        ncell = np.atleast_1d(self.flags).size
        return "Result(%d cell%s, ierr %d, %d failed)" % (
            ncell, "" if ncell == 1 else "s", self.ierr, self.failed().size)
