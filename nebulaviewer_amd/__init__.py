"""nebulaviewer_amd: NebulaViewer's simulation step (isothermal Euler FV + Stromgren ionization + FUV) as
hand-written HIP kernels for MI355X (gfx950), behind the reference's Sim interface.

    from nebulaviewer_amd.sim import Sim
"""
__all__ = ["sim", "scenarios"]
