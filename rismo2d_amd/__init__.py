"""MI355X (gfx950) library for the Rismo2D UVS2D flow-solve hot path.

``rismo2d_amd.rmx`` binds the HIP library (librmx.so, C ABI in include/rmx.h);
``rismo2d_amd.deck`` builds the synthetic meshes and boundary lines it solves on.
"""
