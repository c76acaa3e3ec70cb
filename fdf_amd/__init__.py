"""fdf_amd: gfx950 (MI355X) implementation of PICurv's per-step particle loop.

The product is the C-ABI library fdf_amd/libpicgpu.so (include/picgpu.h);
`fdf_amd.swarm.PicGpu` is the host-side mirror of the reference's entry
points over that ABI, `fdf_amd.workloads` the synthetic benchmark inputs.
"""
__version__ = "0.1.0"
