"""fastcone_amd — MI355X (gfx950) cone-beam CT projector / backprojector.

A from-scratch HIP implementation behind the reference Fastcone interfaces:
  * ``libFastcone.so`` exporting the reference's five ``GPU_CB_*`` C entry points plus additive
    ``fc_*`` device-pointer entry points (declared in include/fastcone.h);
  * ``fastcone_amd.fastcone``: the reference's Python call surface (project, backproject, fbp,
    rampfilter);
  * ``fastcone_amd.ops``: torch (ROCm) operators, the matched-adjoint autograd Function and FDK.
"""
__version__ = "0.1.0"
