"""CuPyRAM host driver for the MI355X march.

``CuPyRAM(freq, zs, zr, z_ss, rp_ss, cw, z_sb, rp_sb, cb, rhob, attn, rbzb,
**kwargs).run()`` keeps the constructor, keyword defaults, grid rules and the
returned dictionary of cupyram/cupyram.py:45-380.  The host resolves the
per-ray parameters (cupyram.py:460-608), packs the environment tables, and
computes the FP64 Pade coefficients; everything per range step -- range
update, profiles, Galerkin assembly, the np tridiagonal solves and the TL
output -- runs on the GPU inside libcupyram_hip.so with no host round trip.

Deliberate differences from the reference, each a fix of a reference defect:
  * the TL line is computed whether or not ``compute_grids`` is set (the
    reference leaves it at zero with grids off, cupyram.py:306-307);
  * the CP grid is written (outpt.py:50 assigns a local instead);
  * the range-dependence of profiles and bathymetry is decided per ray, not
    from ray 0 (cupyram.py:437-442);
  * 'Proc Time' is wall-clock time of run() (the reference reports CPU
    process time, cupyram.py:288,320);
  * the self-starter's (1-X)^-1 smoothing applies one factor as PyRAM does
    (the reference's fused kernel also sweeps np-1 identity factors);
  * the CP Grid is reshaped and squeezed exactly like the TL Grid (the
    reference leaves it [N_env, 1, nvz, nvr] for one frequency and squeezes
    the wrong axis for one environment, cupyram.py:356-368).

Extensions (keyword arguments the reference does not have):
  precision='c128'|'c64'   arithmetic of the march (FP64 Pade + starter in both)
  variant='auto'|'thread'|'team'   kernel family
  freq_per_ray=True        ``freq`` is an array of one frequency per ray
  dz / dr / rmax as arrays one value per ray: per-ray grids (then the
                           wavelength of the absorbing layer is c0/f per ray)
  steps_per_launch=k       march in launches of k range steps captured in a
                           hipGraph (0: one launch for the whole march)
  compute_cells=True       adds 'TL Cells' (shaped like 'TL Grid'): the TL of
                           the mean of |p|^2/r over each output cell, ranges
                           (i ndr, (i+1) ndr] dr x rows (j ndz, (j+1) ndz] --
                           the coarse TL the learned surrogate is trained on
"""
from __future__ import annotations

import time

import numpy

from . import _abi
from .pade import compute_pade_batch


class CuPyRAM:

    _np_default = 8
    _dzf = 0.1
    _ndr_default = 1
    _ndz_default = 1
    _ns_default = 1
    _lyrw_default = 20
    _id_default = 0

    # -- input normalisation (cupyram.py:55-174) --------------------------------

    @staticmethod
    def _normalize_to_batch(x, batch_size, param_name):
        """List of per-ray arrays -> NaN-padded [batch, ...]; a single array or
        scalar is tiled over the batch."""
        if isinstance(x, list):
            if len(x) != batch_size:
                raise ValueError(f"{param_name}: list length {len(x)} != batch_size {batch_size}")
            arrays = [numpy.asarray(item) for item in x]
            shapes = {a.shape for a in arrays}
            if len(shapes) == 1:
                return numpy.stack(arrays, axis=0)
            if len({a.ndim for a in arrays}) != 1:
                raise ValueError(f"{param_name}: per-ray arrays differ in dimensionality")
            ndim = arrays[0].ndim
            full = tuple(max(a.shape[i] for a in arrays) for i in range(ndim))
            dtype = arrays[0].dtype
            if numpy.issubdtype(dtype, numpy.integer) or dtype == bool:
                dtype = numpy.float64
            out = numpy.full((batch_size,) + full, numpy.nan, dtype=dtype)
            for i, a in enumerate(arrays):
                out[(i,) + tuple(slice(0, s) for s in a.shape)] = a
            return out
        arr = numpy.asarray(x)
        if arr.ndim == 0:
            return numpy.full(batch_size, arr)
        return numpy.tile(arr, (batch_size,) + (1,) * arr.ndim)

    @staticmethod
    def _get_valid_slice(arr):
        """Strip NaN padding: 1-D up to the first NaN; 2-D drops all-NaN rows
        and columns; returns (array, length)."""
        arr = numpy.asarray(arr)
        if arr.ndim == 1:
            nan = numpy.isnan(arr)
            if not nan.any():
                return arr, len(arr)
            n = int(numpy.argmax(nan))
            return (arr[:n], n) if n > 0 else (arr, len(arr))
        if arr.ndim == 2:
            rows = int(numpy.sum(~numpy.all(numpy.isnan(arr), axis=1)))
            cols = int(numpy.sum(~numpy.all(numpy.isnan(arr), axis=0)))
            if rows == arr.shape[0] and cols == arr.shape[1]:
                return arr, arr.shape[0]
            return arr[:rows, :cols], rows
        flat = arr.reshape(arr.shape[0], -1)
        ok = ~numpy.any(numpy.isnan(flat), axis=1)
        if ok.all():
            return arr, len(arr)
        n = int(numpy.argmax(~ok))
        return (arr[:n], n) if n > 0 else (arr, len(arr))

    # -- construction (cupyram.py:176-271) ---------------------------------------

    def __init__(self, freq, zs, zr, z_ss, rp_ss, cw, z_sb, rp_sb, cb, rhob, attn, rbzb, **kwargs):
        self._batch_size = int(kwargs.get("batch_size", 1))
        self._compute_grids = bool(kwargs.get("compute_grids", True))
        self._compute_cells = bool(kwargs.get("compute_cells", False))
        self._max_workers = kwargs.get("max_workers", 8)
        self._precision = {"c64": 64, "complex64": 64, 64: 64, "c128": 128, "complex128": 128,
                           128: 128}[kwargs.get("precision", "c128")]
        variant = kwargs.get("variant", "auto")
        if variant not in _abi.VARIANTS:
            raise ValueError(f"variant must be one of {sorted(_abi.VARIANTS)}")
        self._variant = variant
        self._team_shape = kwargs.get("team_shape", None)
        self._steps_per_launch = int(kwargs.get("steps_per_launch", 0))
        self._freq_per_ray = bool(kwargs.get("freq_per_ray", False))
        if numpy.isscalar(freq):
            self._freqs = numpy.array([freq], dtype=numpy.float64)
        else:
            self._freqs = numpy.array(freq, dtype=numpy.float64).flatten()
        if self._freq_per_ray:
            if self._freqs.size == 1:
                self._freqs = numpy.full(self._batch_size, self._freqs[0])
            if self._freqs.size != self._batch_size:
                raise ValueError("freq_per_ray needs one frequency per ray")
            self._n_freq = 1
        else:
            self._n_freq = len(self._freqs)
        if numpy.any(~(self._freqs > 0)):
            raise ValueError("frequencies must be positive")
        self._n_env = self._batch_size
        self._total_batch = self._n_env * self._n_freq
        # zs may be one value per ray (extension, like the per-ray dz/dr arrays)
        self._zs_e = self._per_env(zs, self._batch_size)
        self._zs = float(zs) if numpy.ndim(zs) == 0 else self._zs_e
        self._zr = float(zr)
        z_ss = self._normalize_to_batch(z_ss, self._batch_size, "z_ss")
        rp_ss = self._normalize_to_batch(rp_ss, self._batch_size, "rp_ss")
        cw = self._normalize_to_batch(cw, self._batch_size, "cw")
        z_sb = self._normalize_to_batch(z_sb, self._batch_size, "z_sb")
        rp_sb = self._normalize_to_batch(rp_sb, self._batch_size, "rp_sb")
        cb = self._normalize_to_batch(cb, self._batch_size, "cb")
        rhob = self._normalize_to_batch(rhob, self._batch_size, "rhob")
        attn = self._normalize_to_batch(attn, self._batch_size, "attn")
        rbzb = self._normalize_to_batch(rbzb, self._batch_size, "rbzb")
        self.check_inputs(z_ss, rp_ss, cw, z_sb, rp_sb, cb, rhob, attn, rbzb)
        self.get_params(**kwargs)

    @staticmethod
    def _valid_len(A):
        """Per-ray valid length of NaN-padded rows [B, n]: the _get_valid_slice
        rule for 1-D arrays (up to the first NaN; all of it if the first
        element is NaN)."""
        nan = numpy.isnan(A)
        first = nan.argmax(axis=1)
        return numpy.where(nan.any(axis=1) & (first > 0), first, A.shape[1])

    @staticmethod
    def _valid_shape(A):
        """Per-ray (rows, cols) of NaN-padded blocks [B, r, c]: the
        _get_valid_slice rule for 2-D arrays (counts of rows and columns that
        are not all NaN; the valid block is the leading one)."""
        nan = numpy.isnan(A)
        return (~nan.all(axis=2)).sum(axis=1), (~nan.all(axis=1)).sum(axis=1)

    def check_inputs(self, z_ss, rp_ss, cw, z_sb, rp_sb, cb, rhob, attn, rbzb):
        """Validate every ray and keep its valid (un-padded) tables
        (cupyram.py:382-458).  Works on the whole NaN-padded batch at once;
        the first failing ray raises the reference's message."""
        self._status_ok = True
        B = self._batch_size
        f64 = lambda a: numpy.asarray(a, numpy.float64)  # noqa: E731
        T = {"z_ss": f64(z_ss).reshape(B, -1), "rp_ss": f64(rp_ss).reshape(B, -1),
             "z_sb": f64(z_sb).reshape(B, -1), "rp_sb": f64(rp_sb).reshape(B, -1),
             "rbzb": f64(rbzb).reshape(B, -1, 2)}
        for name, arr in (("cw", cw), ("cb", cb), ("rhob", rhob), ("attn", attn)):
            a = f64(arr)
            T[name] = a.reshape(B, a.shape[1], -1) if a.ndim >= 2 else a.reshape(B, -1, 1)
        L = {k: self._valid_len(T[k]) for k in ("z_ss", "rp_ss", "z_sb", "rp_sb")}
        S = {k: self._valid_shape(T[k]) for k in ("cw", "cb", "rhob", "attn", "rbzb")}
        ar = numpy.arange(B)
        zfirst = T["z_ss"][:, 0]
        zlast = T["z_ss"][ar, L["z_ss"] - 1]
        bt_rows = S["rbzb"][0]
        bt_z = numpy.where(numpy.arange(T["rbzb"].shape[1])[None, :] < bt_rows[:, None], T["rbzb"][:, :, 1], -numpy.inf)
        bt_max = bt_z.max(axis=1)

        def decreasing(A, n):
            d = numpy.diff(A, axis=1)
            ok = numpy.arange(d.shape[1])[None, :] < (n - 1)[:, None]
            return numpy.any(ok & (d < 0), axis=1)

        checks = [
            (~((zfirst <= self._zs_e) & (self._zs_e <= zlast)),
             lambda b: f"Ray {b}: Source depth {self._zs_e[b]}m outside range [{zfirst[b]}, {zlast[b]}]m"),
            (~((zfirst <= self._zr) & (self._zr <= zlast)),
             lambda b: f"Ray {b}: Receiver depth {self._zr}m outside range [{zfirst[b]}, {zlast[b]}]m"),
            ((S["cw"][0] != L["z_ss"]) | (S["cw"][1] != L["rp_ss"]),
             lambda b: f"Ray {b}: z_ss ({L['z_ss'][b]}), rp_ss ({L['rp_ss'][b]}), "
                       f"cw ({S['cw'][0][b]}, {S['cw'][1][b]}) inconsistent"),
        ]
        for name in ("cb", "rhob", "attn"):
            checks.append(((S[name][0] != L["z_sb"]) | (S[name][1] != L["rp_sb"]),
                           lambda b, name=name: f"Ray {b}: z_sb ({L['z_sb'][b]}), rp_sb ({L['rp_sb'][b]}), "
                                                f"{name} ({S[name][0][b]}, {S[name][1][b]}) inconsistent"))
        checks.append((bt_max > zlast, lambda b: f"Ray {b}: Max bathy ({bt_max[b]}m) > max SSP depth ({zlast[b]}m)"))
        checks.append((decreasing(T["z_ss"], L["z_ss"]) | decreasing(T["z_sb"], L["z_sb"]),
                       lambda b: f"Ray {b}: profile depths must be non-decreasing"))
        bad = numpy.zeros(B, bool)
        for m, _ in checks:
            bad |= m
        if bad.any():
            b = int(numpy.argmax(bad))
            for m, msg in checks:
                if m[b]:
                    raise ValueError(msg(b))
        self._tables, self._lens, self._shapes = T, L, S
        self._zss_last = zlast
        self._env = [{"z_ss": T["z_ss"][b, :L["z_ss"][b]], "rp_ss": T["rp_ss"][b, :L["rp_ss"][b]],
                      "cw": T["cw"][b, :S["cw"][0][b], :S["cw"][1][b]],
                      "z_sb": T["z_sb"][b, :L["z_sb"][b]], "rp_sb": T["rp_sb"][b, :L["rp_sb"][b]],
                      "cb": T["cb"][b, :S["cb"][0][b], :S["cb"][1][b]],
                      "rhob": T["rhob"][b, :S["rhob"][0][b], :S["rhob"][1][b]],
                      "attn": T["attn"][b, :S["attn"][0][b], :S["attn"][1][b]],
                      "rbzb": T["rbzb"][b, :bt_rows[b], :S["rbzb"][1][b]]} for b in range(B)]
        e0 = self._env[0]
        self.rd_ss = e0["rp_ss"].size > 1
        self.rd_sb = e0["rp_sb"].size > 1
        self.rd_bt = e0["rbzb"].shape[0] > 1

    def _masked(self, name, col=None):
        """Valid entries of table `name` for every ray, NaN elsewhere (or the
        given column of the blocks)."""
        A = self._tables[name]
        if name in self._lens:
            return numpy.where(numpy.arange(A.shape[1])[None, :] < self._lens[name][:, None], A, numpy.nan)
        rows, cols = self._shapes[name]
        ok = (numpy.arange(A.shape[1])[None, :, None] < rows[:, None, None]) & \
             (numpy.arange(A.shape[2])[None, None, :] < cols[:, None, None])
        M = numpy.where(ok, A, numpy.nan)
        return M if col is None else M[:, :, col]

    def get_params(self, **kwargs):
        """Defaults of cupyram.py:460-526; per-ray arrays are an extension."""
        self._np = int(kwargs.get("np", CuPyRAM._np_default))
        if not 1 <= self._np <= _abi.MAX_NP:
            raise ValueError(f"np must be in 1..{_abi.MAX_NP}")
        self._benchmark_propagation = kwargs.get("benchmark_propagation", False)
        if self._benchmark_propagation:
            self.propagation_time = 0.0
        if "c0" in kwargs:
            self._c0 = kwargs["c0"]
            self._c0_array = numpy.full(self._batch_size, float(kwargs["c0"]))
        else:
            with numpy.errstate(invalid="ignore"):
                self._c0_array = numpy.nanmean(self._masked("cw", 0), axis=1)
            self._c0 = float(numpy.mean(self._c0_array))
        max_freq = float(numpy.max(self._freqs))
        self._lambda = self._c0 / max_freq
        self._per_ray_grid = any(numpy.ndim(kwargs.get(k, 0.0)) > 0 for k in ("dz", "dr", "rmax", "ndr", "ndz"))
        self._dr = kwargs.get("dr", self._np * 1500 / max_freq)
        self._dz = kwargs.get("dz", CuPyRAM._dzf * 1500 / max_freq)
        # output decimation; per-ray arrays are an extension of the reference's scalars
        self._ndr_e = self._per_env(kwargs.get("ndr", CuPyRAM._ndr_default), self._batch_size).astype(numpy.int64)
        self._ndz_e = self._per_env(kwargs.get("ndz", CuPyRAM._ndz_default), self._batch_size).astype(numpy.int64)
        if numpy.any(self._ndr_e < 1) or numpy.any(self._ndz_e < 1):
            raise ValueError("ndr and ndz must be >= 1")
        self._ndr = int(self._ndr_e[0]) if numpy.ndim(kwargs.get("ndr", 1)) == 0 else self._ndr_e
        self._ndz = int(self._ndz_e[0]) if numpy.ndim(kwargs.get("ndz", 1)) == 0 else self._ndz_e
        self._zmplt_given = "zmplt" in kwargs
        self._bt_zmax = numpy.nanmax(self._masked("rbzb", 1), axis=1)
        self._zmplt = kwargs["zmplt"] if "zmplt" in kwargs else float(numpy.max(self._bt_zmax))
        rmax_default = max(float(numpy.nanmax(self._masked("rp_ss"))), float(numpy.nanmax(self._masked("rp_sb"))),
                           float(numpy.nanmax(self._masked("rbzb", 0))))
        self._rmax = kwargs.get("rmax", rmax_default)
        self._rmax_input = self._rmax
        self._ns = int(kwargs.get("ns", CuPyRAM._ns_default))
        self._rs = kwargs.get("rs", numpy.asarray(self._rmax) + numpy.asarray(self._dr))
        self._lyrw = kwargs.get("lyrw", CuPyRAM._lyrw_default)
        self._id = kwargs.get("id", CuPyRAM._id_default)
        self.proc_time = None

    # -- setup (cupyram.py:528-726) ----------------------------------------------

    def _per_env(self, value, n):
        v = numpy.asarray(value, numpy.float64)
        if v.ndim == 0:
            return numpy.full(n, float(v))
        if v.size != n:
            raise ValueError("per-ray parameter arrays must have one value per ray")
        return v.ravel()

    def setup(self):
        """Resolve per-ray grids, pack the environment and the Pade sets.

        It depends on the constructor's arguments only, so the work is done
        once: run() calls it again (as the reference's run() does) and returns
        at once when it has been called before, for instance by a thread that
        prepares the next batch while the GPU marches this one."""
        if getattr(self, "_setup_done", False):
            return
        ne, nf, nc = self._n_env, self._n_freq, self._total_batch
        dz_e = self._per_env(self._dz, ne)
        dr_e = self._per_env(self._dr, ne)
        rmax_e = self._per_env(self._rmax_input, ne)
        rs_e = self._per_env(self._rs, ne)
        if self._freq_per_ray:
            f_env = self._freqs
        else:
            f_env = numpy.full(ne, float(numpy.max(self._freqs)))
        if self._per_ray_grid:
            lam_e = self._c0_array / f_env
        else:
            lam_e = numpy.full(ne, self._lambda)
        # seabed depths are relative to the last water-profile depth (cupyram.py:580-587)
        zsb_abs = self._masked("z_sb") + self._zss_last[:, None]
        zmax_sb = numpy.nanmax(zsb_abs, axis=1)
        if self._per_ray_grid:
            zmax_e = zmax_sb + self._lyrw * lam_e
        else:
            zmax_e = numpy.full(ne, float(numpy.max(zmax_sb)) + self._lyrw * self._lambda)
        nz_e = numpy.floor(zmax_e / dz_e).astype(numpy.int64) - 1
        if self._per_ray_grid and not self._zmplt_given:
            zmplt_e = self._bt_zmax.copy()
        else:
            zmplt_e = self._per_env(self._zmplt, ne)
        nzplt_e = numpy.floor(zmplt_e / dz_e).astype(numpy.int64)
        self._zmplt_e = zmplt_e
        ndr_e, ndz_e = self._ndr_e, self._ndz_e
        nvr_e = numpy.floor(rmax_e / (dr_e * ndr_e)).astype(numpy.int64)
        rmax_trunc = nvr_e * dr_e * ndr_e
        nsteps_e = numpy.round(rmax_trunc / dr_e).astype(numpy.int64) - 1
        nvz_e = numpy.floor(nzplt_e / ndz_e).astype(numpy.int64)
        if numpy.any(nz_e < 3):
            raise ValueError("depth grid too coarse (nz < 3)")
        if numpy.any(nvz_e * ndz_e > nz_e + 1):
            raise ValueError("zmplt lies below the computational grid")
        # environment tables, extended bathymetry (cupyram.py:536-568), packed
        # table by table: every ray's valid block of one table, then the next
        T, Ls, Sh = self._tables, self._lens, self._shapes
        ar = numpy.arange(ne)
        bt_rows = Sh["rbzb"][0]
        last = bt_rows - 1
        ext = T["rbzb"][:, :, :2]
        ext = numpy.concatenate([ext, numpy.full((ne, 2, 2), numpy.nan)], axis=1)
        grow = T["rbzb"][ar, last, 0] < rmax_e          # extend to rmax with the last depth
        ext[ar[grow], bt_rows[grow], 0] = rmax_e[grow]
        ext[ar[grow], bt_rows[grow], 1] = T["rbzb"][ar[grow], last[grow], 1]
        n_bt = bt_rows + grow
        one = n_bt == 1                                 # a single point: a flat second one 1 m on
        ext[ar[one], 1, 0] = ext[ar[one], 0, 0] + 1.0
        ext[ar[one], 1, 1] = ext[ar[one], 0, 1]
        n_bt = numpy.where(one, 2, n_bt)
        iz0_e = numpy.clip(numpy.floor(ext[:, 0, 1] / dz_e).astype(numpy.int64), 1, nz_e - 1)
        chunks, envs, off = [], numpy.zeros(ne, _abi.ENV_DTYPE), 0

        def pack(field, A, rows, cols=None):
            """Append every ray's leading rows (x cols) block of A; set offsets."""
            nonlocal off
            if cols is None:
                ok = numpy.arange(A.shape[1])[None, :] < rows[:, None]
                size = rows
            else:
                ok = (numpy.arange(A.shape[1])[None, :, None] < rows[:, None, None]) & \
                     (numpy.arange(A.shape[2])[None, None, :] < cols[:, None, None])
                size = rows * cols
            chunks.append(numpy.ascontiguousarray(A[ok], numpy.float64))
            envs[field] = off + numpy.concatenate([[0], numpy.cumsum(size)[:-1]])
            off += int(numpy.sum(size))

        envs["n_ss"], envs["n_rp_ss"] = Sh["cw"]
        envs["n_sb"], envs["n_rp_sb"] = Sh["cb"]
        envs["n_bt"] = n_bt
        pack("z_ss", T["z_ss"], Ls["z_ss"])
        pack("rp_ss", T["rp_ss"], Ls["rp_ss"])
        pack("cw", T["cw"], *Sh["cw"])
        pack("z_sb", zsb_abs, Ls["z_sb"])
        pack("rp_sb", T["rp_sb"], Ls["rp_sb"])
        for name in ("cb", "rhob", "attn"):
            pack(name, T[name], *Sh[name])
        pack("rbzb", ext, n_bt, numpy.full(ne, 2))
        self._env_data = numpy.concatenate(chunks) if chunks else numpy.zeros(1)
        env_of = numpy.arange(nc) // nf
        if self._freq_per_ray:
            f_calc = self._freqs
        else:
            f_calc = numpy.tile(self._freqs, ne)
        rays = numpy.zeros(nc, _abi.RAY_DTYPE)
        rays["dz"] = dz_e[env_of]
        rays["zmax"] = zmax_e[env_of]
        rays["nz"] = nz_e[env_of]
        rays["zstep"] = zmax_e[env_of] / (nz_e[env_of] + 1)
        rays["dr"] = dr_e[env_of]
        rays["omega"] = 2 * numpy.pi * f_calc
        rays["c0"] = self._c0_array[env_of]
        rays["k0"] = rays["omega"] / rays["c0"]
        rays["lyrw_lambda"] = self._lyrw * lam_e[env_of]
        rays["rs"] = rs_e[env_of]
        rays["zs"] = self._zs_e[env_of]
        rays["zr"] = self._zr
        rays["nsteps"] = nsteps_e[env_of]
        rays["iz0"] = iz0_e[env_of]
        rays["ndr"] = ndr_e[env_of]
        rays["ndz"] = ndz_e[env_of]
        rays["nvr"] = nvr_e[env_of]
        rays["nvz"] = nvz_e[env_of]
        rays["np"] = self._np
        self._rays = rays
        self._envs = envs[env_of]
        # Pade sets per calculation: march (ip=1, ns), starter (ip=2, ns), after rs (ip=1, ns=0)
        pade = numpy.zeros((nc, 3, 2, self._np), numpy.complex128)
        key = numpy.stack([f_calc, rays["c0"], rays["dr"]], axis=1)
        uniq, inv = numpy.unique(key, axis=0, return_inverse=True)
        inv = inv.ravel()
        for s, (ip, ns) in enumerate(((1, self._ns), (2, self._ns), (1, 0))):
            sets = compute_pade_batch(uniq[:, 0], uniq[:, 1], self._np, ns, uniq[:, 2], ip)
            pade[:, s] = sets[inv]
        self._pade = pade
        # attributes of the reference after setup
        self.nz = int(nz_e.max())
        self.nzplt = int(nzplt_e.max())
        self._zmax = float(zmax_e.max()) if not self._per_ray_grid else zmax_e
        if not self._per_ray_grid:
            self._rmax = float(rmax_trunc[0])
        nvr, nvz = int(nvr_e.max()), int(nvz_e.max())
        if self._per_ray_grid:
            self.vr = [numpy.arange(1, n + 1) * d * m for n, d, m in zip(nvr_e, dr_e, ndr_e)]
            self.vz = [numpy.arange(1, n + 1) * d * m for n, d, m in zip(nvz_e, dz_e, ndz_e)]
        else:
            self.vr = numpy.arange(1, nvr + 1) * dr_e[0] * ndr_e[0]
            self.vz = numpy.arange(1, nvz + 1) * dz_e[0] * ndz_e[0]
        self._nvr, self._nvz = nvr, nvz
        self.eta = 1 / (40 * numpy.pi * numpy.log10(numpy.exp(1)))
        self._setup_done = True

    # -- run (cupyram.py:273-380) ------------------------------------------------

    def _make_plan(self, idx, stream=None, shape=None, variant=None):
        cfg = _abi.CprConfig()
        cfg.nrays = len(idx)
        cfg.precision = self._precision
        cfg.variant = _abi.VARIANTS[variant or self._variant]
        cfg.compute_grids = 1 if self._compute_grids else 0
        cfg.nvr_max = int(self._rays["nvr"][idx].max())
        cfg.compute_cells = 1 if self._compute_cells else 0
        cfg.nvz_max = int(self._rays["nvz"][idx].max()) if (self._compute_grids or self._compute_cells) else 0
        shape = shape or self._team_shape
        if shape:
            cfg.team_rows, cfg.team_lanes = shape
        return _abi.Plan(cfg, self._rays[idx], self._envs[idx], self._env_data, self._pade[idx], stream)

    def _buckets(self):
        """Group the calculations by kernel shape: one plan per register-team
        shape (rows per lane x lanes per ray), one for the thread march.  A
        batch whose rays share a grid is a single bucket, as in the reference;
        uneven batches are no longer padded to the deepest ray."""
        nc = self._total_batch
        if self._variant == "thread" or self._team_shape:
            return [(numpy.arange(nc), self._team_shape, self._variant)]
        from .autotune import tuned_shape
        nz = self._rays["nz"]
        table = {}
        shapes = []
        for z in nz:
            z = int(z)
            if z not in table:
                table[z] = tuned_shape(self._precision, z) or _abi.team_shape(self._precision, z)
            shapes.append(table[z])
        if self._variant == "team" and any(sh == (0, 0) for sh in shapes):
            raise ValueError("some rays are too deep for the register-resident team march; use variant='auto'")
        groups = {}
        for c, sh in enumerate(shapes):
            groups.setdefault(sh, []).append(c)
        # deepest buckets first: their rays have the longest dependent chain
        # per step (fewest rays, largest columns), so they bound the job's end
        # unless they start at once; the wide shallow buckets fill around them
        # Within a bucket the rays go longest first (most range steps): a
        # block of teams (or a wavefront of thread-march lanes) holds its CU
        # slots until its longest ray ends, so rays of similar length share
        # blocks, and the long ones start at once instead of in the tail.
        nsteps = self._rays["nsteps"]
        out = []
        for sh, ids in sorted(groups.items(), key=lambda kv: (-int(nz[kv[1]].max()), -len(kv[1]))):
            ids = numpy.array(ids)
            ids = ids[numpy.argsort(-nsteps[ids], kind="stable")]
            if sh == (0, 0):
                out.append((ids, None, "thread"))
            else:
                out.append((ids, sh, "team"))
        return out

    def run(self):
        """March every ray; returns the reference's result dictionary."""
        t0 = time.perf_counter()
        self.setup()
        nc = self._total_batch
        buckets = self._buckets()
        tll = numpy.zeros((nc, self._nvr))
        cpl = numpy.zeros((nc, self._nvr), numpy.complex128)
        tlg = numpy.zeros((nc, self._nvz, self._nvr)) if self._compute_grids else None
        cpg = numpy.zeros((nc, self._nvz, self._nvr), numpy.complex128) if self._compute_grids else None
        tlc = None
        if self._compute_cells:
            # NaN marks the cells past a ray's own grid; without such rays every cell is written
            ragged = numpy.any(self._rays["nvz"] < self._nvz) or numpy.any(self._rays["nvr"] < self._nvr)
            dims = (nc, self._nvz, self._nvr)
            tlc = numpy.full(dims, numpy.nan) if ragged else numpy.empty(dims)
        plans = []
        try:
            for idx, shape, variant in buckets:
                plans.append((idx, self._make_plan(idx, shape=shape, variant=variant)))
            for _, plan in plans:
                plan.selfs()
            for idx, plan in plans:
                nmax = int(self._rays["nsteps"][idx].max())
                if self._steps_per_launch > 0:
                    plan.march_graph(self._steps_per_launch)
                else:
                    plan.march(0, nmax)
            self.march_ms, self.march_launches = 0.0, 0
            self.plan_variants = []
            for idx, plan in plans:
                plan.sync()
                ms, n = plan.march_time()
                self.march_ms = max(self.march_ms, ms)
                self.march_launches += n
                a, b, g, h = plan.read_outputs(grids=self._compute_grids)
                nv = a.shape[1]
                tll[idx, :nv], cpl[idx, :nv] = a, b
                if self._compute_grids:
                    tlg[idx, :g.shape[1], :nv], cpg[idx, :g.shape[1], :nv] = g, h
                if self._compute_cells:
                    cells = plan.read_cells()
                    # one copy per distinct (depth bins, range bins) of the bucket
                    dims = numpy.stack([self._rays["nvz"][idx], self._rays["nvr"][idx]], axis=1)
                    for nz_c, nr_c in numpy.unique(dims, axis=0):
                        k = numpy.nonzero((dims[:, 0] == nz_c) & (dims[:, 1] == nr_c))[0]
                        tlc[idx[k], :nz_c, :nr_c] = cells[k, :nz_c, :nr_c]
                self.plan_variants.append((len(idx), plan.describe(0)))
            self.plan_variant = self.plan_variants[0][1]
        finally:
            for _, plan in plans:
                plan.close()
        self.proc_time = time.perf_counter() - t0
        if self._benchmark_propagation:
            self.propagation_time = self.march_ms / 1e3
        nvr_c = self._rays["nvr"]
        if numpy.any(nvr_c < self._nvr):
            for c in numpy.nonzero(nvr_c < self._nvr)[0]:
                tll[c, nvr_c[c]:] = numpy.nan
                cpl[c, nvr_c[c]:] = numpy.nan
        self.tll, self.cpl, self.tlg, self.cpg = tll, cpl, tlg, cpg
        ne, nf = self._n_env, self._n_freq
        self.tlc = None
        if self._compute_cells:
            # same [env, freq, depth, range] reshaping and squeezing as the TL Grid
            tlc = tlc.reshape(ne, nf, tlc.shape[1], tlc.shape[2])
            self.tlc = tlc[0, 0] if (nf == 1 and ne == 1) else (tlc[:, 0] if nf == 1 else (tlc[0] if ne == 1 else tlc))
        self.tll = self.tll.reshape(ne, nf, -1)
        self.cpl = self.cpl.reshape(ne, nf, -1)
        if self._compute_grids:
            self.tlg = self.tlg.reshape(ne, nf, self.tlg.shape[1], self.tlg.shape[2])
            self.cpg = self.cpg.reshape(ne, nf, self.cpg.shape[1], self.cpg.shape[2])
        if nf == 1 and ne == 1:
            self.tll, self.cpl = self.tll[0, 0], self.cpl[0, 0]
            if self._compute_grids:
                self.tlg, self.cpg = self.tlg[0, 0], self.cpg[0, 0]
        elif nf == 1:
            self.tll, self.cpl = self.tll[:, 0, :], self.cpl[:, 0, :]
            if self._compute_grids:
                self.tlg, self.cpg = self.tlg[:, 0], self.cpg[:, 0]
        elif ne == 1:
            self.tll, self.cpl = self.tll[0], self.cpl[0]
            if self._compute_grids:
                self.tlg, self.cpg = self.tlg[0], self.cpg[0]
        return {"ID": self._id,
                "Proc Time": self.proc_time,
                "Ranges": self.vr,
                "Depths": self.vz,
                "TL Grid": self.tlg,
                "TL Line": self.tll,
                "CP Grid": self.cpg,
                "CP Line": self.cpl,
                "c0": self._c0,
                **({"TL Cells": self.tlc} if self._compute_cells else {})}
