"""One simulation resident on the GPU: the Python face of an ``fdtd_sim`` handle.

Host arrays use the reference's layout, float32 C-order ``(nx, ny, nz)`` (z fastest); the
device stores x fastest, so every transfer transposes. Volumes are named as the reference's
``Fields`` names them (``Dx``, ``fu_Bx``, ``f_w_Ex``, ``P[0].Ex``, ``P_prev[0].Ex``...).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, Optional, Sequence, Tuple

import numpy

from . import _native
from ._native import check, dptr, fptr, iptr, u8ptr
from .dispersion import Susceptibility
from .grid import Grid
from .pml import PML

_AXES = ("x", "y", "z")


def to_device_layout(array: numpy.ndarray) -> numpy.ndarray:
    """C-order (nx, ny, nz) -> contiguous (nz, ny, nx), i.e. x fastest."""
    return numpy.ascontiguousarray(numpy.asarray(array, dtype=numpy.float32).transpose(2, 1, 0))


def from_device_layout(buffer: numpy.ndarray) -> numpy.ndarray:
    return numpy.ascontiguousarray(buffer.transpose(2, 1, 0))


class HipSimulation:
    """A handle on device-resident fields, layer table, materials, poles, sources, monitors."""

    def __init__(self, grid: Grid, pml: Optional[PML] = None, subnormals: str = "keep"):
        """``subnormals``: the build to step with, "keep" (the default) or "flush"."""
        self.lib = _native.load(subnormals=subnormals)
        self.subnormals = subnormals
        self.grid = grid
        self.pml = pml if (pml is not None and pml.is_active) else None
        cfg = _native.FdtdConfig()
        for a in range(3):
            cfg.n[a] = grid.shape[a]
            cfg.invariant[a] = 1 if grid.is_invariant(a) else 0
        cfg.dtdx = numpy.float32(grid.dt / grid.dx)
        cfg.pml = 1 if self.pml is not None else 0
        cfg.dt = grid.dt
        handle = ctypes.c_void_p()
        self._check(self.lib.fdtd_sim_create(ctypes.byref(cfg), ctypes.byref(handle)), "fdtd_sim_create")
        self.handle = handle
        self.poles = []
        self.planes = []
        for axis, fold in enumerate(getattr(grid, "folds", (None, None, None))):
            if fold is not None:
                phase, kind, far_row = fold
                self.set_fold(axis, kind, phase, far_row or 0)
        if self.pml is not None:
            for (half, axis), (kms, sinv, kps) in self.pml.vectors.items():
                member = self.pml.members[(half, axis)].astype(numpy.uint8)
                self._check(self.lib.fdtd_set_layer(self.handle, half, axis, fptr(kms), fptr(sinv),
                                              fptr(kps), u8ptr(member)), "fdtd_set_layer")

    def _check(self, rc: int, what: str) -> None:
        check(rc, what, self.lib)

    # --- lifetime ----------------------------------------------------------------------------
    #: the live ``BatchedRun`` stepping this simulation (a weak reference), or None
    _batch = None

    def batch(self):
        """The live batch this simulation is a member of, or None."""
        ref = self._batch
        b = ref() if ref is not None else None
        return b if b is not None and getattr(b, "handle", None) is not None else None

    def close(self):
        """Free the device state. Refused while a live batch steps the simulation: its
        descriptors point into this memory, so close the ``BatchedRun`` first."""
        if self.batch() is not None:
            raise RuntimeError("this simulation is a member of a live BatchedRun: close the "
                               "batch before the simulation")
        self._release()

    def _release(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            # a member of a batch the engine still holds is released with that batch
            self.lib.fdtd_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        out = ctypes.c_uint64()
        self._check(self.lib.fdtd_sim_bytes(self.handle, ctypes.byref(out)), "fdtd_sim_bytes")
        return int(out.value)

    # --- materials ---------------------------------------------------------------------------
    def set_inverse_epsilon(self, component: str, volume: numpy.ndarray):
        dev = to_device_layout(volume)
        self._check(self.lib.fdtd_set_inverse_epsilon(self.handle, _AXES.index(component[1]),
                                                fptr(dev)), "fdtd_set_inverse_epsilon")

    def set_boundary(self, axis: int, kind: str):
        """One axis's boundary: "periodic" or "metallic" (``fdtd_set_boundary``)."""
        kinds = ("periodic", "metallic")
        if kind not in kinds:
            raise ValueError(f"boundary kind must be one of {kinds}, got {kind!r}")
        self._check(self.lib.fdtd_set_boundary(self.handle, int(axis), kinds.index(kind)),
              "fdtd_set_boundary")

    def set_fold(self, axis: int, kind: int, phase: int = 1, far_row: int = 0):
        """A mirror fold of one axis (``fdtd_set_fold``): kind 1 over metallic boundaries, 2
        over periodic ones (``far_row``: the row the far ghost images), 0 none."""
        self._check(self.lib.fdtd_set_fold(self.handle, int(axis), int(kind), int(phase),
                                           int(far_row)), "fdtd_set_fold")

    def set_conductivity(self, component: str, volume: Optional[numpy.ndarray],
                         cells: Optional[numpy.ndarray] = None):
        """MEEP's D or B conductivity of one component ("Dx".."Bz"): sigma as a float32 volume
        in the reference layout, or None / all zero for none (``fdtd_set_conductivity``).
        ``cells`` (bool, reference layout) confines it to the cells whose chunk keeps it
        (``fdtd_set_conductive_cells``); None keeps it everywhere."""
        if len(component) != 2 or component[0] not in "DB" or component[1] not in _AXES:
            raise ValueError(f"a conductivity belongs to Dx..Dz or Bx..Bz, not {component!r}")
        dev = None if volume is None else to_device_layout(
            numpy.broadcast_to(numpy.asarray(volume, dtype=numpy.float32), self.grid.shape))
        self._check(self.lib.fdtd_set_conductivity(self.handle, 0 if component[0] == "D" else 1,
                                             _AXES.index(component[1]),
                                             fptr(dev) if dev is not None else None),
              "fdtd_set_conductivity")
        if cells is not None and dev is not None and numpy.any(dev):
            mask = numpy.ascontiguousarray(numpy.broadcast_to(
                numpy.asarray(cells, dtype=bool), self.grid.shape).astype(numpy.uint8)
                .transpose(2, 1, 0))  # x fastest
            self._check(self.lib.fdtd_set_conductive_cells(self.handle,
                                                     0 if component[0] == "D" else 1,
                                                     _AXES.index(component[1]), u8ptr(mask)),
                  "fdtd_set_conductive_cells")

    def pole_row_fraction(self) -> float:
        """Fraction of the grid's cells inside the per-row ranges where pole work runs."""
        active, rows = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.fdtd_pole_rows(self.handle, ctypes.byref(active), ctypes.byref(rows)),
              "fdtd_pole_rows")
        return active.value / rows.value if rows.value else 0.0

    def work_split(self, vec: int = 4) -> Dict[str, object]:
        """The fused launches' split for runs of ``vec`` cells: the interior box (stepped by the
        lossless forms) and the run counts of the interior and of the shell around it."""
        out = numpy.zeros(8, dtype=numpy.int32)
        self._check(self.lib.fdtd_work_split(self.handle, int(vec), iptr(out)), "fdtd_work_split")
        return {"box": tuple(int(v) for v in out[:6]), "interior_runs": int(out[6]),
                "shell_runs": int(out[7])}

    def inverse_epsilon_bytes(self) -> Tuple[int, int, int]:
        """Bytes per cell the D pair reads for each component's inverse epsilon (1, 2 or 4)."""
        out = numpy.zeros(3, dtype=numpy.int32)
        self._check(self.lib.fdtd_inverse_epsilon_bytes(self.handle, iptr(out)),
              "fdtd_inverse_epsilon_bytes")
        return int(out[0]), int(out[1]), int(out[2])

    def set_offdiagonal(self, row: str, partner: str, volume: numpy.ndarray):
        dev = to_device_layout(volume)
        self._check(self.lib.fdtd_set_offdiagonal(self.handle, _AXES.index(row[1]),
                                            _AXES.index(partner[1]), fptr(dev)),
              "fdtd_set_offdiagonal")

    def add_pole(self, term: Optional[Susceptibility], sigma: Dict[str, object],
                 constants: Optional[Sequence[float]] = None) -> Tuple[bool, bool, bool]:
        """Install one pole; ``sigma[Ex..]`` a volume or a scalar. ``constants`` (the four
        float32 words of ``realnum_constants``) replaces the words formed from ``term``, as a
        lifted run carries them. Returns which components it drives."""
        words = constants if constants is not None else term.realnum_constants(self.grid.dt)
        consts = numpy.array(words, dtype=numpy.float32)
        if consts.shape != (4,):
            raise ValueError(f"a pole has 4 constant words, got {consts.shape}")
        keep = []
        ptrs = (_native._FP * 3)()
        scalars = numpy.zeros(3, dtype=numpy.float32)
        driven = numpy.zeros(3, dtype=numpy.int32)
        for a, c in enumerate(_AXES):
            value = sigma.get("E" + c, 0.0)
            if numpy.ndim(value) == 0:
                scalars[a] = numpy.float32(value)
                driven[a] = 1 if float(value) != 0.0 else 0
                ptrs[a] = ctypes.cast(None, _native._FP)
            else:
                vol = to_device_layout(value)
                keep.append(vol)
                driven[a] = 1 if numpy.any(vol != 0.0) else 0
                ptrs[a] = fptr(vol)
        index = numpy.zeros(1, dtype=numpy.int32)
        self._check(self.lib.fdtd_add_pole(self.handle, fptr(consts), ptrs, fptr(scalars), iptr(driven),
                                     iptr(index)), "fdtd_add_pole")
        self.poles.append(tuple(bool(v) for v in driven))
        return self.poles[-1]

    def add_source(self, magnetic: bool, component: int, lo: Sequence[int], cnt: Sequence[int],
                   amp_re: numpy.ndarray, amp_im: numpy.ndarray, current: numpy.ndarray,
                   dt: float, integrated: bool = False):
        """A deposit over box [lo, lo+cnt); amplitudes C-order (cnt), current (nsteps, 2) f64.
        ``integrated``: an electric polarization source, ``current`` its dipole table."""
        re = numpy.ascontiguousarray(numpy.asarray(amp_re, dtype=numpy.float64)
                                     .reshape(tuple(cnt)).transpose(2, 1, 0))
        im = numpy.ascontiguousarray(numpy.asarray(amp_im, dtype=numpy.float64)
                                     .reshape(tuple(cnt)).transpose(2, 1, 0))
        cur = numpy.ascontiguousarray(current, dtype=numpy.float64).reshape(-1, 2)
        lo_a = numpy.asarray(lo, dtype=numpy.int32)
        cnt_a = numpy.asarray(cnt, dtype=numpy.int32)
        field = (3 if magnetic else 2) if integrated else (1 if magnetic else 0)
        self._check(self.lib.fdtd_add_source(self.handle, field, int(component),
                                       iptr(lo_a), iptr(cnt_a), dptr(re), dptr(im), dptr(cur),
                                       int(cur.shape[0]), float(dt)), "fdtd_add_source")

    def set_source_currents(self, index: int, current: numpy.ndarray):
        """Replace source ``index``'s current table (``fdtd_set_source_currents``): (nsteps, 2)
        float64 from step 0, e.g. a longer one for a run past the planned length."""
        cur = numpy.ascontiguousarray(current, dtype=numpy.float64).reshape(-1, 2)
        self._check(self.lib.fdtd_set_source_currents(self.handle, int(index), dptr(cur),
                                                      int(cur.shape[0])),
                    "fdtd_set_source_currents")

    # --- state -------------------------------------------------------------------------------
    def upload(self, name: str, array: numpy.ndarray):
        dev = to_device_layout(array)
        self._check(self.lib.fdtd_upload(self.handle, name.encode(), fptr(dev), dev.size),
              f"fdtd_upload({name})")

    def download(self, name: str) -> numpy.ndarray:
        nx, ny, nz = self.grid.shape
        buf = numpy.empty((nz, ny, nx), dtype=numpy.float32)
        self._check(self.lib.fdtd_download(self.handle, name.encode(), fptr(buf), buf.size),
              f"fdtd_download({name})")
        return from_device_layout(buf)

    @property
    def step_count(self) -> int:
        out = numpy.zeros(1, dtype=numpy.int32)
        self._check(self.lib.fdtd_get_step(self.handle, iptr(out)), "fdtd_get_step")
        return int(out[0])

    def set_step(self, step: int, h_separated: bool):
        self._check(self.lib.fdtd_set_step(self.handle, int(step), 1 if h_separated else 0),
              "fdtd_set_step")

    def prime_e(self):
        self._check(self.lib.fdtd_prime_e(self.handle), "fdtd_prime_e")

    # --- stepping ----------------------------------------------------------------------------
    def step_h(self):
        self._check(self.lib.fdtd_step_h(self.handle), "fdtd_step_h")

    def step_d(self):
        self._check(self.lib.fdtd_step_d(self.handle), "fdtd_step_d")

    def run(self, steps: int):
        self._check(self.lib.fdtd_run(self.handle, int(steps)), "fdtd_run")

    def synchronize(self):
        self._check(self.lib.fdtd_synchronize(self.handle), "fdtd_synchronize")

    def set_config(self, config: Optional["LaunchConfig"]):
        """This simulation's own launch configuration; None: follow the process defaults."""
        words = config.words() if config is not None else None
        self._check(self.lib.fdtd_sim_set_config(self.handle, iptr(words) if words is not None
                                                 else None), "fdtd_sim_set_config")

    def config(self) -> "LaunchConfig":
        """The launch configuration in force for this simulation."""
        out = numpy.zeros(LaunchConfig.WORDS, dtype=numpy.int32)
        self._check(self.lib.fdtd_sim_get_config(self.handle, iptr(out)), "fdtd_sim_get_config")
        return LaunchConfig.from_words(out)

    def graph_stats(self) -> Tuple[int, int]:
        """(graphs captured, replays after their capture) over this simulation's life."""
        out = (ctypes.c_int64 * 2)()
        self._check(self.lib.fdtd_graph_stats(self.handle, out), "fdtd_graph_stats")
        return int(out[0]), int(out[1])

    def flush(self):
        """Transform every plane's buffered samples now (``fdtd_flush``), in stream order."""
        self._check(self.lib.fdtd_flush(self.handle), "fdtd_flush")

    def run_timed(self, steps: int) -> Tuple[float, float, float]:
        """(total ms, mean H-pair ms, mean D-pair ms) of ``steps`` steps, HIP events per launch."""
        out = numpy.zeros(3, dtype=numpy.float32)
        self._check(self.lib.fdtd_run_timed(self.handle, int(steps), fptr(out)), "fdtd_run_timed")
        return float(out[0]), float(out[1]), float(out[2])

    def time_pairs(self, repeats: int) -> Tuple[float, float]:
        out = numpy.zeros(2, dtype=numpy.float32)
        self._check(self.lib.fdtd_time_pairs(self.handle, int(repeats), fptr(out)), "fdtd_time_pairs")
        return float(out[0]), float(out[1])

    # --- flux --------------------------------------------------------------------------------
    def add_flux_plane(self, normal: int, lo: Sequence[int], cnt: Sequence[int],
                       omegas: numpy.ndarray, scale: float, decimation: int = 1) -> int:
        om = numpy.ascontiguousarray(omegas, dtype=numpy.float64)
        lo_a = numpy.asarray(lo, dtype=numpy.int32)
        cnt_a = numpy.asarray(cnt, dtype=numpy.int32)
        out = numpy.zeros(1, dtype=numpy.int32)
        self._check(self.lib.fdtd_add_flux_plane(self.handle, int(normal), iptr(lo_a), iptr(cnt_a),
                                           int(om.size), dptr(om), float(scale), int(decimation),
                                           iptr(out)), "fdtd_add_flux_plane")
        self.planes.append((int(normal), tuple(int(v) for v in lo), tuple(int(v) for v in cnt),
                            int(om.size), 4))
        return int(out[0])

    def add_dft_region(self, lo: Sequence[int], cnt: Sequence[int], components: Sequence[str],
                       omegas: numpy.ndarray, scale: float, decimation: int = 1,
                       yee_grid: bool = False) -> int:
        """A DFT field region over the cells [lo, lo+cnt): up to 4 components, cell-centred or
        at their own Yee sites (``fdtd_add_dft_region``). Read back with ``flux_read``."""
        names = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
        if not 1 <= len(components) <= 4 or any(c not in names for c in components):
            raise ValueError(f"a DFT region takes 1 to 4 of {names}, got {tuple(components)}")
        om = numpy.ascontiguousarray(omegas, dtype=numpy.float64)
        lo_a = numpy.asarray(lo, dtype=numpy.int32)
        cnt_a = numpy.asarray(cnt, dtype=numpy.int32)
        codes = numpy.asarray([names.index(c) for c in components], dtype=numpy.int32)
        out = numpy.zeros(1, dtype=numpy.int32)
        self._check(self.lib.fdtd_add_dft_region(self.handle, iptr(lo_a), iptr(cnt_a), int(codes.size),
                                           iptr(codes), 1 if yee_grid else 0, int(om.size),
                                           dptr(om), float(scale), int(decimation), iptr(out)),
              "fdtd_add_dft_region")
        self.planes.append((-1, tuple(int(v) for v in lo), tuple(int(v) for v in cnt),
                            int(om.size), int(codes.size)))
        return int(out[0])

    def flux_read(self, plane: int) -> numpy.ndarray:
        """complex128 transforms [4 (E1, E2, H1, H2), or a region's components][nf][cnt...] in
        C-order over the box."""
        normal, lo, cnt, nf, nslot = self.planes[plane]
        cells = int(numpy.prod(cnt))
        buf = numpy.empty(2 * nslot * nf * cells, dtype=numpy.float64)
        self._check(self.lib.fdtd_flux_read(self.handle, int(plane), dptr(buf), buf.size),
              "fdtd_flux_read")
        data = buf.view(numpy.complex128).reshape(nslot, nf, cnt[2], cnt[1], cnt[0])
        return numpy.ascontiguousarray(data.transpose(0, 1, 4, 3, 2))

    def flux_spectrum(self, plane: int, cell_weights: Optional[numpy.ndarray] = None
                      ) -> numpy.ndarray:
        """sum over the plane of w * Re(E1 conj(H2) - E2 conj(H1)) per frequency, reduced on the
        device (``fdtd_flux_spectrum``); ``cell_weights`` C-order over the box, or None for 1."""
        normal, lo, cnt, nf, nslot = self.planes[plane]
        out = numpy.empty(nf, dtype=numpy.float64)
        w = None
        if cell_weights is not None:
            w = numpy.ascontiguousarray(numpy.broadcast_to(
                numpy.asarray(cell_weights, dtype=numpy.float64), tuple(cnt)).transpose(2, 1, 0))
        self._check(self.lib.fdtd_flux_spectrum(self.handle, int(plane),
                                                dptr(w) if w is not None else None, dptr(out),
                                                out.size), "fdtd_flux_spectrum")
        return out

    def flux_write(self, plane: int, transforms: numpy.ndarray) -> None:
        """Replace a plane's transforms: complex [4][nf][cnt...] as ``flux_read`` returns them."""
        normal, lo, cnt, nf, nslot = self.planes[plane]
        data = numpy.asarray(transforms)
        if data.shape != (nslot, nf) + tuple(cnt):
            raise ValueError(f"plane {plane} takes transforms of shape "
                             f"{(nslot, nf) + tuple(cnt)}, got {data.shape}")
        buf = numpy.ascontiguousarray(data.astype(numpy.complex128).transpose(0, 1, 4, 3, 2))
        flat = buf.view(numpy.float64).reshape(-1)
        self._check(self.lib.fdtd_flux_write(self.handle, int(plane), dptr(flat), flat.size),
              "fdtd_flux_write")

    def flux_scale(self, plane: int, scale: complex) -> None:
        """Multiply a plane's transforms by ``scale``."""
        z = complex(scale)
        self._check(self.lib.fdtd_flux_scale(self.handle, int(plane), z.real, z.imag), "fdtd_flux_scale")


def device_info() -> str:
    lib = _native.load()
    buf = ctypes.create_string_buffer(256)
    check(lib.fdtd_device_info(buf, 256), "fdtd_device_info")
    return buf.value.decode()


def copy_probe(nbytes: int, repeats: int = 20, mode: int = 0) -> float:
    """Mean ms of a device copy of ``nbytes`` (profile calibration; mode as fdtd_copy_probe)."""
    out = numpy.zeros(1, dtype=numpy.float32)
    check(_native.load().fdtd_copy_probe(int(nbytes), int(repeats), int(mode), fptr(out)),
          "fdtd_copy_probe")
    return float(out[0])


@dataclass(frozen=True)
class LaunchConfig:
    """How one simulation's (or batch's) launches run: per fused pair (H, D) the shape
    ``(cells per thread, workgroups per CU or 0, tile order, non-temporal stores, shell cells
    per thread, halo tile rows)``, graph replay on or off, and the flux transform (1, 2 or 4
    MFMA tiles per wave, 0 the vector kernel). Halo tile rows (the D pair only, one cell a
    thread): 0 for runs of cells, or 2, 4, 8 for the box's rows over LDS-staged tiles of
    ``256 / rows`` x ``rows`` cells. ``fdtd_sim_set_config`` / ``fdtd_batch_set_config``
    words."""
    h: Tuple[int, int, int, bool, int, int] = (1, 8, 0, False, 1, 0)
    d: Tuple[int, int, int, bool, int, int] = (1, 8, 0, False, 1, 0)
    graphs: bool = True
    flux_mfma: int = 1

    #: words per pair's shape, and in all
    SHAPE_WORDS = 6
    WORDS = 14

    def words(self) -> numpy.ndarray:
        out = [int(v) for v in _full(self.h)] + [int(v) for v in _full(self.d)]
        return numpy.array(out + [1 if self.graphs else 0, int(self.flux_mfma)], dtype=numpy.int32)

    @classmethod
    def from_words(cls, w) -> "LaunchConfig":
        w = [int(v) for v in w]
        h = (w[0], w[1], w[2], bool(w[3]), w[4], w[5])
        d = (w[6], w[7], w[8], bool(w[9]), w[10], w[11])
        return cls(h, d, bool(w[12]), w[13])

    @classmethod
    def of_shapes(cls, shapes, shell=(1, 1), graphs: bool = True, flux_mfma: int = 1,
                  tile: int = 0):
        """From (H shape, D shape) 4-tuples as the autotuner returns them (a D shape may carry
        its halo tile rows as a fifth entry; ``tile`` otherwise)."""
        h = tuple(shapes[0][:4]) + (int(shell[0]), 0)
        dt = int(shapes[1][4]) if len(shapes[1]) > 4 else int(tile)
        d = tuple(shapes[1][:4]) + (int(shell[1]), dt)
        return cls(h, d, graphs, flux_mfma)


def _full(shape) -> Tuple:
    """A pair's shape with its halo tile rows (0 when a 5-entry shape leaves them out)."""
    shape = tuple(shape)
    return shape + (0,) * (6 - len(shape))


def link_parts(re: "HipSimulation", im: "HipSimulation", kinds: Sequence[int],
               phases: Sequence[Tuple[float, float]]):
    """Make two simulations the real and imaginary parts of one complex storage
    (``fdtd_link_parts``): per axis the Bloch connection kind (0 none, 1 copy, 2 negate,
    3 a product) and the phase's parts, rounded here to float32 (MEEP's complex<realnum>)."""
    k = numpy.asarray([int(v) for v in kinds], dtype=numpy.int32)
    ph = numpy.asarray([numpy.float32(x) for pair in phases for x in pair], dtype=numpy.float32)
    check(re.lib.fdtd_link_parts(re.handle, im.handle, iptr(k),
                                 ph.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
          "fdtd_link_parts", re.lib)


class ComplexSimulation:
    """Complex field storage: two linked ``HipSimulation`` parts, the real and the imaginary
    (reference ``Fields(force_complex_fields=True)``, part_products.py:1-40). The parts step
    only together, as a batch of two whose one launch per half-step serves both; any other
    operation (a read, a table write) ends that batch first, after its work has drained."""

    def __init__(self, re: HipSimulation, im: HipSimulation):
        self.sims = (re, im)
        self.grid, self.pml, self.lib = re.grid, re.pml, re.lib
        self._run = None

    @property
    def handle(self):
        return self.sims[0].handle

    def _stepper(self):
        if self._run is None:
            import types

            from .batch import BatchedRun

            self._run = BatchedRun([types.SimpleNamespace(sim=s, grid=s.grid)
                                    for s in self.sims])
        return self._run

    def _settle(self):
        if self._run is not None:
            try:
                self._run.synchronize()
            finally:
                self._run.close()
                self._run = None

    def step_h(self):
        self._stepper().step_h()

    def step_d(self):
        self._stepper().step_d()

    def run(self, steps: int):
        self._stepper().run(int(steps))

    def synchronize(self):
        if self._run is not None:
            self._run.synchronize()
        else:
            for s in self.sims:
                s.synchronize()

    @property
    def step_count(self) -> int:
        return self.sims[0].step_count

    def download(self, name: str, part: int = 0) -> numpy.ndarray:
        self._settle()
        return self.sims[part].download(name)

    def download_complex(self, name: str) -> numpy.ndarray:
        """complex64: the real part's words plus i times the imaginary part's."""
        self._settle()
        re, im = (s.download(name) for s in self.sims)
        out = numpy.empty(re.shape, dtype=numpy.complex64)
        out.real, out.imag = re, im
        return out

    # --- flux and DFT regions: one region on each part, read as one complex transform ---------
    @property
    def planes(self):
        return self.sims[0].planes

    def add_flux_plane(self, *args, **kw) -> int:
        self._settle()
        ids = [s.add_flux_plane(*args, **kw) for s in self.sims]
        assert ids[0] == ids[1]
        return ids[0]

    def add_dft_region(self, *args, **kw) -> int:
        self._settle()
        ids = [s.add_dft_region(*args, **kw) for s in self.sims]
        assert ids[0] == ids[1]
        return ids[0]

    def flux_read(self, plane: int) -> numpy.ndarray:
        """The transforms of the complex field: the real part's plus i times the imaginary
        part's (the transform is linear, and each part's is complex itself)."""
        self._settle()
        re, im = (s.flux_read(plane) for s in self.sims)
        return re + 1j * im

    def flux_spectrum(self, plane: int, cell_weights: Optional[numpy.ndarray] = None
                      ) -> numpy.ndarray:
        """sum over the plane of w * Re(E1 conj(H2) - E2 conj(H1)) of the complex field's
        transforms, on the host: the two parts' cross terms need both planes' cells."""
        from .dft import flux_spectrum

        return flux_spectrum(self.flux_read(plane), 1.0, 1.0, cell_weights)

    def flux_write(self, plane: int, transforms: numpy.ndarray) -> None:
        """The whole transform into the real part and zeros into the imaginary part: their
        combination is the transform written, and later samples add to both linearly."""
        self._settle()
        data = numpy.asarray(transforms)
        self.sims[0].flux_write(plane, data)
        self.sims[1].flux_write(plane, numpy.zeros_like(data))

    def flux_scale(self, plane: int, scale: complex) -> None:
        self._settle()
        for s in self.sims:
            s.flux_scale(plane, scale)

    def flush(self):
        self._settle()
        for s in self.sims:
            s.flush()

    def set_source_currents(self, index: int, currents: numpy.ndarray):
        """The same current table on both parts (each keeps its own amplitudes)."""
        self._settle()
        for s in self.sims:
            s.set_source_currents(index, currents)

    def close(self):
        self._settle()
        for s in self.sims:
            s.close()


def set_launch(vec: int, blocks_per_cu: int, xcd_remap: int, nt_stores: bool = False):
    """Process-wide launch shape: cells per thread, workgroups per CU (0 = one per tile), tile
    order over the XCDs (0 plain, 1 contiguous eighths, sb >= 2 super-blocks of sb tiles),
    non-temporal output stores."""
    check(_native.load().fdtd_set_launch(int(vec), int(blocks_per_cu), int(xcd_remap),
                                         1 if nt_stores else 0), "fdtd_set_launch")


def get_launch() -> Tuple[int, int, int, bool]:
    """The H pair's launch shape (``set_launch`` sets both pairs alike)."""
    return get_launch_pair(0)


def set_launch_pair(pair: int, vec: int, blocks_per_cu: int, xcd_remap: int,
                    nt_stores: bool = False):
    """One fused pair's launch shape: pair 0 = H pair, 1 = D pair."""
    check(_native.load().fdtd_set_launch_pair(int(pair), int(vec), int(blocks_per_cu),
                                              int(xcd_remap), 1 if nt_stores else 0),
          "fdtd_set_launch_pair")


def set_launch_shell(pair: int, vec: int):
    """Cells per thread of one pair's shell kernel (the layer forms outside the interior box)."""
    check(_native.load().fdtd_set_launch_shell(int(pair), int(vec)), "fdtd_set_launch_shell")


def set_launch_tile(rows: int):
    """The D pair's box rows over LDS-staged halo tiles of ``rows`` rows (2, 4, 8), or 0."""
    check(_native.load().fdtd_set_launch_tile(int(rows)), "fdtd_set_launch_tile")


def get_launch_tile() -> int:
    out = numpy.zeros(1, dtype=numpy.int32)
    check(_native.load().fdtd_get_launch_tile(iptr(out)), "fdtd_get_launch_tile")
    return int(out[0])


def get_launch_shell(pair: int) -> int:
    out = numpy.zeros(1, dtype=numpy.int32)
    check(_native.load().fdtd_get_launch_shell(int(pair), iptr(out)), "fdtd_get_launch_shell")
    return int(out[0])


def get_launch_pair(pair: int) -> Tuple[int, int, int, bool]:
    out = numpy.zeros(4, dtype=numpy.int32)
    check(_native.load().fdtd_get_launch_pair(int(pair), iptr(out)), "fdtd_get_launch_pair")
    return int(out[0]), int(out[1]), int(out[2]), bool(out[3])


def set_flux_transform(mfma: bool, cell_tiles: int = 1):
    """Run the float64 flux transform on the matrix cores (default; ``cell_tiles`` 16-cell tiles
    per wave: 1, 2 or 4) or as vector FMAs."""
    check(_native.load().fdtd_set_flux_transform(int(cell_tiles) if mfma else 0),
          "fdtd_set_flux_transform")


def set_graphs(enable: bool):
    """Replay captured 128-step hipGraphs in ``run`` and batched runs (default on)."""
    check(_native.load().fdtd_set_graphs(1 if enable else 0), "fdtd_set_graphs")


def device_memory() -> Tuple[int, int]:
    """(free, total) bytes of device memory on the current GPU."""
    free, total = ctypes.c_uint64(), ctypes.c_uint64()
    check(_native.load().fdtd_device_memory(ctypes.byref(free), ctypes.byref(total)),
          "fdtd_device_memory")
    return int(free.value), int(total.value)


#: flux samples buffered per plane before a transform pass (kRing in the engine)
RING = 128


def footprint_bytes(shape: Sequence[int], pml: bool, driven_poles: Sequence[Tuple[int, int]] = (),
                    offdiagonal_volumes: int = 0,
                    planes: Sequence[Tuple[int, int]] = (), source_steps: int = 0,
                    source_points: int = 0, conductive_components: int = 0) -> int:
    """Device bytes a simulation allocates (an upper bound on ``HipSimulation.device_bytes``).

    ``driven_poles``: per pole (driven components, sigma volumes); ``planes``: per flux plane
    (cells, frequencies). Mirrors the engine's allocations: 12 float32 volumes (D, B, E and
    inverse epsilon), 15 more under an absorber (H, fu_B, fu_D, f_w_E, f_w_H), two per driven
    pole component plus its sigma volume, the ring, phase and float64 accumulators per plane,
    and the fused launches' work lists.
    """
    nx, ny, nz = (int(v) for v in shape)
    npad = ((nx + 3) // 4) * 4 * ny * nz
    vol = 4 * npad
    total = 4096 + vol * (12 + (15 if pml else 0))
    total += 16 * 2 * (npad // max(ny * nz, 1) + ny + nz)  # the six layer coefficient lines
    # inverse epsilon as table indices (at most 2 bytes a cell) and tables (<= 65536 words)
    total += 3 * (2 * npad + 4 * min(65536, nx * ny * nz) + 512)
    if driven_poles:
        total += ny * nz + 256  # the rows where pole work runs
    for comps, sig in driven_poles:
        total += vol * (2 * int(comps) + int(sig))
    total += vol * int(offdiagonal_volumes)
    # per conductive component: condfac, condinv and, under an absorber, f_cond
    total += vol * (3 if pml else 2) * int(conductive_components)
    for cells, nf in planes:
        row = ((int(cells) + 15) // 16) * 16
        total += 8 * nf + 4 * RING * 4 * row + 8 * RING + 8 * RING * 4 * nf + 64 * nf * cells
        total += 512 * (len(planes) + 1)
    total += 16 * source_points + 16 * (source_steps + 1) + 256
    # the work split: shell lists for runs of 1, 2 and 4 cells (at most every run) and the
    # work items of kThreads runs each, per width
    total += 4 * (npad + npad // 2 + npad // 4) + 16 * 3 * (npad // 256 + 2) + 64
    return int(total)


def stream_probe(nread: int, nwrite: int, n: int, stagger: int = 0, repeats: int = 10) -> float:
    """Mean ms of one launch touching ``nread`` + ``nwrite`` float streams of ``n`` cells."""
    out = numpy.zeros(1, dtype=numpy.float32)
    check(_native.load().fdtd_stream_probe(int(nread), int(nwrite), int(n), int(stagger),
                                           int(repeats), fptr(out)), "fdtd_stream_probe")
    return float(out[0])
