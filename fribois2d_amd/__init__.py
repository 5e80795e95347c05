"""FRiBoIS2D fractional-step fluid step on MI355X (gfx950).

* ``fribois2d_amd.gpu``    -- ctypes binding of libfribo_gpu.so (the product)
* ``fribois2d_amd.cases``  -- grids, markers, initial state and phini inputs
* ``fribois2d_amd.build``  -- in-tree hipcc / amdflang build
* ``fribois2d_amd/fortran`` -- iso_c_binding interface for the Fortran driver
"""
__all__ = ["gpu", "cases", "build"]
