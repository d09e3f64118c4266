"""Python drop-in for the reference's src/d3_cffi.py: class D3Calculator over libd3.so.

Same constructor, methods, defaults and return values as the reference (src/d3_cffi.py:51-129):
float32 arrays in and out, energy in eV, forces (N, 3) in eV/Angstrom, stress (3, 3) in eV/A^3.
The library is loaded from ``<this dir>/../build/libd3.so`` as in the reference.  cffi is used
when importable (the reference's binding); otherwise the ctypes bindings of d3_lib.py, which
declare the same C signatures.  Extension: ``atm_cutoff`` (bohr) enables the three-body term
through init_d3_handle_with_atm, which the reference's cdef does not declare.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

try:  # the reference binds through cffi (src/d3_cffi.py:1-49)
    import cffi  # type: ignore

    _HAVE_CFFI = True
except ImportError:  # cffi is not installed in every environment: use ctypes
    _HAVE_CFFI = False

import d3_lib

_lib = d3_lib.load()

if _HAVE_CFFI:
    ffi = cffi.FFI()
    ffi.cdef("""
    typedef void d3_handle_t;
    d3_handle_t *init_d3_handle(uint16_t *elements, uint64_t length_elements, uint64_t max_length,
                                float cutoff_radius, float coordination_number_cutoff,
                                int damping_type, int functional_type);
    d3_handle_t *init_d3_handle_with_atm(uint16_t *elements, uint64_t length_elements, uint64_t max_length,
                                         float cutoff_radius, float coordination_number_cutoff, float atm_cutoff,
                                         int damping_type, int functional_type);
    void set_atoms(d3_handle_t *handle, float *coords, uint16_t *elements, uint64_t length);
    void set_cell(d3_handle_t *handle, float cell[3][3]);
    void free_d3_handle(d3_handle_t *handle);
    void clear_d3_handle(d3_handle_t *handle);
    uint16_t compute_dispersion_energy_from_handle_status(d3_handle_t *handle, float *energy, float *force,
                                                          float *stress);
    """)
    lib = ffi.dlopen(d3_lib.LIB_PATH)

    def _u16(a):
        return ffi.cast("uint16_t*", a.ctypes.data)

    def _f32(a):
        return ffi.cast("float*", a.ctypes.data)

    def _cell(a):
        return ffi.cast("float (*)[3]", a.ctypes.data)
else:
    ffi = None
    lib = _lib

    def _u16(a):
        return a.ctypes.data_as(C.POINTER(C.c_uint16))

    def _f32(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    _cell = _f32


class D3Calculator:
    def __init__(self, elements: list[int], max_length: int, cutoff_radius: float = 46.475800,
                 cn_cutoff_radius: float = 46.475800, damping_type: int = 0, functional_type: int = 0,
                 atm_cutoff: float | None = None):
        """Initialize a D3 calculator object"""
        elements_array = np.ascontiguousarray(elements, dtype=np.uint16)
        self._elements_keepalive = elements_array
        if atm_cutoff is None:
            self.handle = lib.init_d3_handle(_u16(elements_array), len(elements_array), max_length, cutoff_radius,
                                             cn_cutoff_radius, damping_type, functional_type)
        else:
            self.handle = lib.init_d3_handle_with_atm(_u16(elements_array), len(elements_array), max_length,
                                                      cutoff_radius, cn_cutoff_radius, atm_cutoff, damping_type,
                                                      functional_type)
        if not self.handle:
            raise RuntimeError(f"init_d3_handle failed: {d3_lib.last_error()}")
        print("handle initialized")
        self.num_atoms = max_length

    def set_atoms(self, coords: np.ndarray, elements: list[int]):
        """Set atom coordinates and elements"""
        coords_array = np.ascontiguousarray(coords, dtype=np.float32)
        elements_array = np.ascontiguousarray(elements, dtype=np.uint16)
        lib.set_atoms(self.handle, _f32(coords_array), _u16(elements_array), len(elements_array))
        self.num_atoms = len(elements_array)

    def set_cell(self, cell: np.ndarray):
        """Set periodic cell parameters"""
        cell_array = np.ascontiguousarray(cell, dtype=np.float32)
        lib.set_cell(self.handle, _cell(cell_array))

    def compute(self):
        """Compute dispersion energy, forces, and stress"""
        num_atoms = self.num_atoms
        if num_atoms <= 0:
            raise ValueError("Number of atoms must be greater than zero.")
        forces = np.zeros((num_atoms, 3), dtype=np.float32)
        stress = np.zeros(9, dtype=np.float32)
        if ffi is not None:
            energy = ffi.new("float*")
            result = lib.compute_dispersion_energy_from_handle_status(self.handle, energy, _f32(forces),
                                                                      _f32(stress))
            e = float(energy[0])
        else:
            energy = C.c_float(0.0)
            result = lib.compute_dispersion_energy_from_handle_status(self.handle, C.byref(energy), _f32(forces),
                                                                      _f32(stress))
            e = float(energy.value)
        if result == 0b01:
            # the reference reports every failure this way (src/d3_cffi.py:117-118); the detail
            # is in d3x_last_error
            raise RuntimeError(f"Error computing dispersion energy: {d3_lib.last_error() or 'neighbor list overflow'}")
        return e, forces, stress.reshape(3, 3)

    def clear(self):
        """Clear the D3 handle data"""
        lib.clear_d3_handle(self.handle)

    def __del__(self):
        """Clean up when the object is deleted"""
        if getattr(self, "handle", None):
            lib.free_d3_handle(self.handle)
            self.handle = None
