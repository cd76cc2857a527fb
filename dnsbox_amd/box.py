"""Python front-end of one MI355X box, mirroring the reference module API.

The reference exposes argument-less Fortran module procedures over module
globals (rhs.f90, state.f90, dnsboxheun.f90).  ``Box`` keeps those names and
their meaning on an explicit object:

=====================  ==============================================  =========================
Box method             reference routine                                C-ABI entry point
=====================  ==============================================  =========================
``rhs_int_fact(dt)``   ``rhs_int_fact``            rhs.f90:8-16        ``dnsbox_int_fact``
``rhs_nonlinear()``    ``rhs_nonlinear``           rhs.f90:18-135      ``dnsbox_rhs_nonlinear``
``step(n)``            loop body                   dnsboxheun.f90:81-157  ``dnsbox_step``
``state_project()``    ``state_project``           state.f90:120-179   ``dnsbox_project``
``state_dealias()``    ``state_dealias``           state.f90:181-209   ``dnsbox_project``
``stats()``            ``io_save_stats``           io.f90:628-661      ``dnsbox_diagnostics``
``spectrum()``         ``state_spectrum``          state.f90:350-411   ``dnsbox_diagnostics``
``set_state_physical`` ``state_u2uhat``            state.f90:75-81     ``dnsbox_set_state_physical``
``get_state_physical`` ``state_uhat2u``            state.f90:83-96     ``dnsbox_get_state_physical``
``init_random``        ``state_init_rand``         state.f90:211-285   ``dnsbox_init_random``
=====================  ==============================================  =========================

Host arrays use the reference layout: spectral numpy arrays of shape
(Nx/2+1, Ny, Nz) complex128 (C order == Fortran uhat(Nz, Ny, Nx/2+1)), physical
arrays of shape (Nz, Ny, Nx) float64 (== Fortran u(Nx, Ny, Nz)).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Params, check, load


def _dp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _grid(p: Params):
    """(nx, ny, nz) of a parameter set (each 0 means n)."""
    return tuple(int(v) if v > 0 else int(p.n) for v in (p.nx, p.ny, p.nz))


def default_params(**kw) -> Params:
    p = Params()
    check(load().dnsbox_params_default(C.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


class Box:
    """One periodic box on the GPU: [0, 2 pi / alpha_x) x [0, 2 pi / alpha_y) x [0, 2 pi / alpha_z)
    (alpha_* keyword arguments, default 1) on an nx x ny x nz grid (nx, ny, nz keyword arguments,
    each defaulting to n; a non-cubic grid must be one the library lists, grids.hpp)."""

    def __init__(self, n: int = 32, precision: int = 64, **params):
        self.lib = load()
        self.params = default_params(n=n, precision=precision, **params)
        h = C.c_void_p()
        check(self.lib.dnsbox_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self._owned = True
        self.n = n
        self.nx, self.ny, self.nz = _grid(self.params)
        self.precision = precision

    @classmethod
    def _view(cls, handle, params):
        """A borrowed box of a group (destroyed with the group)."""
        b = cls.__new__(cls)
        b.lib = load()
        b.params = params
        b._h = handle
        b._owned = False
        b.n = params.n
        b.nx, b.ny, b.nz = _grid(params)
        b.precision = params.precision
        return b

    def twin(self) -> "Box":
        """A new box with this box's parameters and pruning setting (its own buffers, state and
        clock), kept with this box and closed with it: the checkpointed backward recomputes
        segments on it while this box runs the step adjoints."""
        tw = getattr(self, "_twin", None)
        if tw is None:
            tw = Box.__new__(Box)
            tw.lib = self.lib
            tw.params = type(self.params).from_buffer_copy(self.params)
            h = C.c_void_p()
            check(self.lib.dnsbox_create(C.byref(tw.params), C.byref(h)))
            tw._h, tw._owned, tw.n, tw.precision = h, True, self.n, self.precision
            tw.nx, tw.ny, tw.nz = self.nx, self.ny, self.nz
            self._twin = tw
        tw.set_pruning(getattr(self, "_prune", True))
        return tw

    # -- lifetime -------------------------------------------------------------------
    def close(self):
        tw = getattr(self, "_twin", None)
        if tw is not None:
            self._twin = None
            tw.close()
        if getattr(self, "_h", None) and self._h.value and getattr(self, "_owned", True):
            check(self.lib.dnsbox_destroy(self._h))
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def spec_shape(self):
        return (self.nx // 2 + 1, self.ny, self.nz)

    @property
    def phys_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def dev_shape(self):
        """Shape of a state tensor in the internal layout: (3, nz, ny, nx/2+1)."""
        return (3, self.nz, self.ny, self.nx // 2 + 1)

    def nspec(self) -> int:
        return int(self.lib.dnsbox_nspec(self._h))

    def sync(self):
        check(self.lib.dnsbox_sync(self._h))

    def set_stream(self, stream_ptr: int):
        check(self.lib.dnsbox_set_stream(self._h, C.c_void_p(stream_ptr)))

    # -- state ----------------------------------------------------------------------
    def set_state(self, uhat, vhat, what):
        arrs = [np.ascontiguousarray(a, np.complex128) for a in (uhat, vhat, what)]
        for a in arrs:
            if a.shape != self.spec_shape:
                raise ValueError(f"expected {self.spec_shape}, got {a.shape}")
        check(self.lib.dnsbox_set_state(self._h, *[_dp(a) for a in arrs]))

    def get_state(self):
        out = [np.zeros(self.spec_shape, np.complex128) for _ in range(3)]
        check(self.lib.dnsbox_get_state(self._h, *[_dp(a) for a in out]))
        return out

    def set_state_physical(self, u, v, w):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (u, v, w)]
        for a in arrs:
            if a.shape != self.phys_shape:
                raise ValueError(f"expected {self.phys_shape}, got {a.shape}")
        check(self.lib.dnsbox_set_state_physical(self._h, *[_dp(a) for a in arrs]))

    def get_state_physical(self):
        out = [np.zeros(self.phys_shape, np.float64) for _ in range(3)]
        check(self.lib.dnsbox_get_state_physical(self._h, *[_dp(a) for a in out]))
        return out

    def save_state(self, path):
        """io_save_state (io.f90:100-278): the physical velocity to an HDF5 state file (datasets
        u, v, w, FP64, Fortran shape (Nx, Ny, Nz))."""
        from . import state_io
        state_io.write_state(path, *self.get_state_physical())

    def load_state(self, path, dealias: bool = True, project: bool = True):
        """io_load_state (io.f90:455-626) followed, as the reference program does
        (dnsboxheun.f90:48-55), by state_u2uhat, state_dealias and state_project."""
        from . import state_io
        dims = state_io.state_dims(path)
        if dims != (self.nx, self.ny, self.nz):
            raise ValueError(f"{path} holds a {dims} grid, the box is {self.nx} x {self.ny} x {self.nz}")
        self.set_state_physical(*state_io.read_state(path))
        if dealias:
            self.state_dealias()
        if project:
            self.state_project()

    def init_random(self, seed: int, box: int = 0, k0: float = 4.0, energy: float = 0.5):
        check(self.lib.dnsbox_init_random(self._h, seed, box, k0, energy))

    def state_project(self):
        check(self.lib.dnsbox_project(self._h, 0, 1))

    def state_dealias(self):
        check(self.lib.dnsbox_project(self._h, 1, 0))

    # -- hot path -------------------------------------------------------------------
    def rhs_int_fact(self, dt: float):
        check(self.lib.dnsbox_int_fact(self._h, dt))

    def rhs_nonlinear(self):
        """N(state) -> returns (nonlinuhat, nonlinvhat, nonlinwhat) in the reference layout."""
        check(self.lib.dnsbox_rhs_nonlinear(self._h))
        out = [np.zeros(self.spec_shape, np.complex128) for _ in range(3)]
        check(self.lib.dnsbox_get_nonlin(self._h, *[_dp(a) for a in out]))
        return out

    def step(self, nsteps: int = 1, stepper: int = 1):
        check(self.lib.dnsbox_step(self._h, nsteps, stepper))

    def time(self):
        t = C.c_double()
        dt = C.c_double()
        check(self.lib.dnsbox_get_time(self._h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def set_time(self, t: float, dt: float):
        check(self.lib.dnsbox_set_time(self._h, t, dt))

    # -- diagnostics ------------------------------------------------------------------
    def diagnostics(self):
        diag = np.zeros(_lib.NDIAG)
        espec = np.zeros(self.nspec())
        check(self.lib.dnsbox_diagnostics(self._h, _dp(diag), _dp(espec)))
        return diag, espec

    def stats(self) -> dict:
        diag, _ = self.diagnostics()
        return dict(zip(_lib.DIAG_NAMES, diag.tolist()))

    def spectrum(self):
        return self.diagnostics()[1]

    def time_passes(self, reps: int = 20):
        """Average duration (ms) of each kernel pass of the step, by HIP events."""
        ms = np.zeros(len(_lib.PASS_NAMES))
        check(self.lib.dnsbox_time_passes(self._h, reps, _dp(ms)))
        return dict(zip(_lib.PASS_NAMES, ms.tolist()))

    def pass_bytes(self):
        """Algorithmic HBM bytes of one launch of each step pass as the next step runs it (every
        element the pass must read or write, once), in the order of time_passes."""
        b = np.zeros(len(_lib.PASS_NAMES))
        check(self.lib.dnsbox_pass_bytes(self._h, _dp(b)))
        return dict(zip(_lib.PASS_NAMES, b.tolist()))

    def set_pruning(self, enable: bool = True):
        """Run the step's transforms pruned to the 2/3-rule cube while the state allows it
        (default on; only exact zeros are skipped, the results agree to rounding)."""
        check(self.lib.dnsbox_set_pruning(self._h, int(bool(enable))))
        self._prune = bool(enable)

    def pruning(self) -> bool:
        """Whether the next step runs pruned."""
        v = C.c_int()
        check(self.lib.dnsbox_get_pruning(self._h, C.byref(v)))
        return bool(v.value)

    def last_step_pruned(self) -> bool:
        """Whether the last step_dev / step_dev_record found its buffer confined to the kept cube
        (and so ran the pruned passes)."""
        v = C.c_int()
        check(self.lib.dnsbox_get_last_step_pruned(self._h, C.byref(v)))
        return bool(v.value)

    def set_vjp_cube(self, confined: bool):
        """Guarantee (True) that the states the step / RHS adjoints linearise about lie in the kept
        cube, so that their first passes read the cube only.  The guarantee is checked on the
        device: withdrawing it (False) raises DnsboxError if a state outside the cube was
        differentiated in between (those adjoints are wrong).  Default False."""
        check(self.lib.dnsbox_set_vjp_cube(self._h, int(bool(confined))))

    def state_ptr(self) -> int:
        p = C.c_void_p()
        check(self.lib.dnsbox_state_ptr(self._h, C.byref(p)))
        return p.value

    # -- device tensors (torch, internal layout (3, nz, ny, nx/2+1), context precision) ----
    def _check_tensor(self, t):
        import torch
        want = torch.complex128 if self.precision == 64 else torch.complex64
        shape = self.dev_shape
        if not (t.is_cuda and t.dtype == want and tuple(t.shape) == shape and t.is_contiguous()):
            raise ValueError(f"expected a contiguous cuda {want} tensor of shape {shape}, got "
                             f"{t.dtype} {tuple(t.shape)} on {t.device}")
        return C.c_void_p(t.data_ptr())

    def state_tensor(self):
        """A device-tensor copy of the state, internal layout (3, nz, ny, nx/2+1)."""
        import torch
        f = np.stack([np.transpose(a, (2, 1, 0)) for a in self.get_state()])
        dt = torch.complex128 if self.precision == 64 else torch.complex64
        return torch.from_numpy(np.ascontiguousarray(f)).to(dt).cuda().contiguous()

    def set_state_tensor(self, t):
        """Set the state from a device tensor of the internal layout."""
        a = t.detach().cpu().numpy().astype(np.complex128)
        self.set_state(*[np.transpose(a[i], (2, 1, 0)) for i in range(3)])

    def empty_state(self):
        import torch
        dt = torch.complex128 if self.precision == 64 else torch.complex64
        return torch.zeros(self.dev_shape, dtype=dt, device="cuda")

    def _on_torch_stream(self):
        import torch
        self.set_stream(torch.cuda.current_stream().cuda_stream)

    def rhs_dev(self, u, out):
        """out <- N(u) (rhs_nonlinear on device tensors)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_rhs_dev(self._h, self._check_tensor(u), self._check_tensor(out)))

    def step_dev(self, state, nsteps: int = 1):
        """IF-Heun steps in place on a device tensor (dt from rhs_int_fact or the controller)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_step_dev(self._h, self._check_tensor(state), nsteps))

    def _check_traj(self, traj, nsteps):
        import torch
        want = torch.complex128 if self.precision == 64 else torch.complex64
        shape = self.dev_shape
        if not (traj.is_cuda and traj.dtype == want and traj.is_contiguous() and tuple(traj.shape[1:]) == shape
                and traj.shape[0] >= nsteps):
            raise ValueError(f"expected a contiguous cuda {want} tensor of shape (>= {nsteps}, {shape}), "
                             f"got {traj.dtype} {tuple(traj.shape)}")
        return C.c_void_p(traj.data_ptr())

    def step_dev_record(self, state, nsteps: int, traj, traj_pred=None):
        """step_dev, storing the state after step j + 1 in traj[j] and, if given, that step's
        predicted state u* in traj_pred[j] (contiguous device tensors of shape
        (>= nsteps, 3, nz, ny, nx/2+1) of the context precision)."""
        self._on_torch_stream()
        tp = self._check_traj(traj_pred, nsteps) if traj_pred is not None else None
        check(self.lib.dnsbox_step_dev_record(self._h, self._check_tensor(state), nsteps,
                                              self._check_traj(traj, nsteps), tp))

    def step_vjp_pred_dev(self, u, ustar, lam, out, dt: float):
        """step_vjp_dev with the step's predicted state u* supplied (from step_dev_record)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_step_vjp_pred_dev(self._h, self._check_tensor(u), self._check_tensor(ustar),
                                                self._check_tensor(lam), self._check_tensor(out), dt))

    @property
    def mixed_slot_shape(self):
        """Shape of one record slot of step_dev_record_mixed: 9 fields of the internal layout."""
        return (9,) + self.dev_shape[1:]

    def _check_mixed(self, t, nslots):
        import torch
        want = torch.complex128 if self.precision == 64 else torch.complex64
        shape = self.mixed_slot_shape
        if not (t.is_cuda and t.dtype == want and t.is_contiguous() and tuple(t.shape[-4:]) == shape
                and t.numel() >= nslots * int(np.prod(shape))):
            raise ValueError(f"expected a contiguous cuda {want} tensor of >= {nslots} slots {shape}, "
                             f"got {t.dtype} {tuple(t.shape)}")
        return C.c_void_p(t.data_ptr())

    def step_dev_record_mixed(self, state, nsteps: int, traj, traj_pred, mixed):
        """step_dev_record that also records the x-pass inputs of the two RHS of step j + 1 in fields
        3..8 of the slots mixed[2 j] and mixed[2 j + 1] (a contiguous tensor of >= 2 nsteps slots of
        shape mixed_slot_shape); returns whether the steps ran pruned (the slots then hold the kept
        kx only)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_step_dev_record_mixed(self._h, self._check_tensor(state), nsteps,
                                                    self._check_traj(traj, nsteps), self._check_traj(traj_pred, nsteps),
                                                    self._check_mixed(mixed, 2 * nsteps)))
        return self.last_step_pruned()

    def step_vjp_mixed_dev(self, u, ustar, lam, out, dt: float, mixed_u, mixed_ustar, pruned_record: bool):
        """step_vjp_pred_dev reading the x-pass inputs at u and u* from the slots that
        step_dev_record_mixed wrote for the step (their fields 0..2 are overwritten)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_step_vjp_mixed_dev(self._h, self._check_tensor(u), self._check_tensor(ustar),
                                                 self._check_tensor(lam), self._check_tensor(out), dt,
                                                 self._check_mixed(mixed_u, 1), self._check_mixed(mixed_ustar, 1),
                                                 int(bool(pruned_record))))

    def step_vjp_mixed_seq_dev(self, u0, traj, traj_pred, lam, out, dt: float, mixed, nsteps: int,
                               pruned_record: bool):
        """step_vjp_mixed_dev over the recorded steps nsteps-1 .. 0 in one call, on what
        step_dev_record_mixed(u0, nsteps, traj, traj_pred, mixed) wrote: step j runs at u0 (j = 0)
        or traj[j - 1], with predicted state traj_pred[j] and the slots mixed[2 j], mixed[2 j + 1].
        The result is left in out; lam is overwritten."""
        self._on_torch_stream()
        if nsteps < 1:
            raise ValueError("nsteps must be >= 1")
        self._check_traj(traj_pred, nsteps)
        if nsteps > 1:
            self._check_traj(traj, nsteps - 1)
        self._check_mixed(mixed, 2 * nsteps)
        step = int(np.prod(self.dev_shape)) * traj_pred.element_size()
        slot = int(np.prod(self.mixed_slot_shape)) * mixed.element_size()
        p0, pt, pp, pm = self._check_tensor(u0).value, traj.data_ptr(), traj_pred.data_ptr(), mixed.data_ptr()
        us = (C.c_void_p * nsteps)(p0, *[pt + (j - 1) * step for j in range(1, nsteps)])
        ps = (C.c_void_p * nsteps)(*[pp + j * step for j in range(nsteps)])
        ms = (C.c_void_p * (2 * nsteps))(*[pm + i * slot for i in range(2 * nsteps)])
        check(self.lib.dnsbox_step_vjp_mixed_seq_dev(self._h, us, ps, self._check_tensor(lam), self._check_tensor(out),
                                                     dt, ms, nsteps, int(bool(pruned_record))))

    def rhs_vjp_dev(self, u, lam, out):
        """out <- J_N(u)^T lam (Euclidean inner product of the stored coordinates)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_rhs_vjp_dev(self._h, self._check_tensor(u), self._check_tensor(lam),
                                          self._check_tensor(out)))

    def step_vjp_dev(self, u, lam, out, dt: float):
        """out <- (d u_{n+1}/d u_n)^T lam for one fixed-dt IF-Heun step from u."""
        self._on_torch_stream()
        check(self.lib.dnsbox_step_vjp_dev(self._h, self._check_tensor(u), self._check_tensor(lam),
                                           self._check_tensor(out), dt))

    def rhs_adj_dev(self, u, out, lin_sign: int = 1):
        """rhsAdj (rhsAdj.f90:356-381, phatadj read as phat); lin_sign=-1: descent direction."""
        self._on_torch_stream()
        check(self.lib.dnsbox_rhs_adj_dev(self._h, self._check_tensor(u), self._check_tensor(out),
                                          lin_sign))

    def adj_descent_step_dev(self, u, dtau: float, nsteps: int = 1, lin_sign: int = 1):
        """The adjoint-descent loop body of dnsboxAdj.f90:76-126, nsteps times, in place on a
        device tensor: Heun in fictitious time on rhsAdj, then state_project."""
        self._on_torch_stream()
        check(self.lib.dnsbox_adj_descent_step_dev(self._h, self._check_tensor(u), dtau, nsteps, lin_sign))

    def adj_descent_step(self, dtau: float, nsteps: int = 1, lin_sign: int = 1):
        """adj_descent_step_dev on the box's own state; its clock advances by dtau per step."""
        check(self.lib.dnsbox_adj_descent_step(self._h, dtau, nsteps, lin_sign))


class BoxGroup:
    """Many independent boxes (N in {16, 32, 48, 64, 128}, one precision) stepped together:
    one launch per pass for all boxes, the step captured in a hipGraph."""

    def __init__(self, params):
        """params: list of dicts of Box keyword arguments (n, precision, nu, ...)."""
        arr = (Params * len(params))()
        for i, kw in enumerate(params):
            arr[i] = default_params(**kw)
        self._create(arr)

    def _create(self, arr):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.dnsbox_group_create(arr, len(arr), C.byref(h)))
        self._h = h
        self.boxes = []
        for i in range(len(arr)):
            bh = C.c_void_p()
            check(self.lib.dnsbox_group_box(h, i, C.byref(bh)))
            self.boxes.append(Box._view(bh, arr[i]))

    def last_step_pruned(self) -> bool:
        """Whether the last step_dev / step_dev_record found every buffer in its kept cube."""
        v = C.c_int()
        check(self.lib.dnsbox_group_get_last_step_pruned(self._h, C.byref(v)))
        return bool(v.value)

    def set_vjp_cube(self, confined: bool):
        """Box.set_vjp_cube for every box of the group."""
        check(self.lib.dnsbox_group_set_vjp_cube(self._h, int(bool(confined))))

    def twin(self) -> "BoxGroup":
        """A new group with this group's box parameters (own buffers, states and clocks), kept
        with this group and closed with it: the checkpointed backward recomputes segments on it
        while this group runs the step adjoints."""
        tw = getattr(self, "_twin", None)
        if tw is None:
            arr = (Params * len(self.boxes))()
            for i, b in enumerate(self.boxes):
                arr[i] = Params.from_buffer_copy(b.params)
            tw = BoxGroup.__new__(BoxGroup)
            tw._create(arr)
            self._twin = tw
        return tw

    def __len__(self):
        return len(self.boxes)

    def __getitem__(self, i):
        return self.boxes[i]

    def step(self, nsteps: int = 1):
        """nsteps IF-Heun steps of every box's own state (hipGraph replay)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_group_step(self._h, nsteps))

    def sync(self):
        check(self.lib.dnsbox_group_sync(self._h))

    def set_stream(self, stream_ptr: int):
        check(self.lib.dnsbox_group_set_stream(self._h, C.c_void_p(stream_ptr)))

    def _on_torch_stream(self):
        import torch
        if torch.cuda.is_available():
            self.set_stream(torch.cuda.current_stream().cuda_stream)

    def empty_states(self):
        return [b.empty_state() for b in self.boxes]

    def _ptrs(self, tensors):
        if len(tensors) != len(self.boxes):
            raise ValueError(f"expected {len(self.boxes)} tensors, got {len(tensors)}")
        arr = (C.c_void_p * len(tensors))()
        for i, (b, t) in enumerate(zip(self.boxes, tensors)):
            arr[i] = b._check_tensor(t).value
        return arr

    def step_dev(self, states, nsteps: int = 1):
        """IF-Heun steps in place on one device tensor per box (each box's own dt)."""
        self._on_torch_stream()
        check(self.lib.dnsbox_group_step_dev(self._h, self._ptrs(states), nsteps))

    def step_vjp_dev(self, us, lams, outs, dts):
        """outs[i] <- (d u_{n+1}/d u_n)^T lams[i] for one fixed-dt step of box i from us[i]."""
        self._on_torch_stream()
        d = (C.c_double * len(self.boxes))(*[float(x) for x in dts])
        check(self.lib.dnsbox_group_step_vjp_dev(self._h, self._ptrs(us), self._ptrs(lams),
                                                 self._ptrs(outs), d))

    def _traj_ptrs(self, trajs, nsteps):
        if len(trajs) != len(self.boxes):
            raise ValueError(f"expected {len(self.boxes)} tensors, got {len(trajs)}")
        arr = (C.c_void_p * len(trajs))()
        for i, (b, t) in enumerate(zip(self.boxes, trajs)):
            arr[i] = b._check_traj(t, nsteps).value
        return arr

    def step_dev_record(self, states, nsteps: int, trajs, traj_preds=None):
        """step_dev, copying box i's state after step j + 1 to trajs[i][j] and, if given, storing
        that step's predicted state u* in traj_preds[i][j] (per box a contiguous device tensor of
        shape (>= nsteps, 3, N, N, N/2+1) of the group precision)."""
        self._on_torch_stream()
        tp = self._traj_ptrs(traj_preds, nsteps) if traj_preds is not None else None
        check(self.lib.dnsbox_group_step_dev_record(self._h, self._ptrs(states), nsteps,
                                                    self._traj_ptrs(trajs, nsteps), tp))

    def _mixed_ptrs(self, mixeds, nslots):
        if len(mixeds) != len(self.boxes):
            raise ValueError(f"expected {len(self.boxes)} tensors, got {len(mixeds)}")
        arr = (C.c_void_p * len(mixeds))()
        for i, (b, t) in enumerate(zip(self.boxes, mixeds)):
            arr[i] = b._check_mixed(t, nslots).value
        return arr

    def step_dev_record_mixed(self, states, nsteps: int, trajs, traj_preds, mixeds):
        """step_dev_record that also records the x-pass inputs of box i's step j + 1 in the slots
        mixeds[i][2 j] and mixeds[i][2 j + 1] (Box.step_dev_record_mixed per box); returns whether
        the steps ran pruned."""
        self._on_torch_stream()
        check(self.lib.dnsbox_group_step_dev_record_mixed(self._h, self._ptrs(states), nsteps,
                                                          self._traj_ptrs(trajs, nsteps),
                                                          self._traj_ptrs(traj_preds, nsteps),
                                                          self._mixed_ptrs(mixeds, 2 * nsteps)))
        return self.last_step_pruned()

    def step_vjp_mixed_dev(self, us, ustars, lams, outs, dts, mixeds, pruned_record: bool):
        """step_vjp_pred_dev reading the recorded x-pass inputs: mixeds[i] is box i's slot pair of
        the step (two slots, the RHS at u then at u*)."""
        self._on_torch_stream()
        d = (C.c_double * len(self.boxes))(*[float(x) for x in dts])
        check(self.lib.dnsbox_group_step_vjp_mixed_dev(self._h, self._ptrs(us), self._ptrs(ustars), self._ptrs(lams),
                                                       self._ptrs(outs), d, self._mixed_ptrs(mixeds, 2),
                                                       int(bool(pruned_record))))

    def step_vjp_pred_dev(self, us, ustars, lams, outs, dts):
        """step_vjp_dev with each box's predicted state u* supplied (from step_dev_record)."""
        self._on_torch_stream()
        d = (C.c_double * len(self.boxes))(*[float(x) for x in dts])
        check(self.lib.dnsbox_group_step_vjp_pred_dev(self._h, self._ptrs(us), self._ptrs(ustars),
                                                      self._ptrs(lams), self._ptrs(outs), d))

    def close(self):
        tw = getattr(self, "_twin", None)
        if tw is not None:
            self._twin = None
            tw.close()
        if getattr(self, "_h", None) and self._h.value:
            for b in self.boxes:
                b._h = C.c_void_p()
            check(self.lib.dnsbox_group_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
