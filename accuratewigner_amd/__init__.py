"""accuratewigner_amd: MI355X-native split-operator propagators for phase-space quantum
dynamics (Wigner-Moyal, Wigner-Bloch, von Neumann, rho-Bloch, Schrodinger-Wigner), a drop-in
for the reference's classes of the same names.

Compute runs in libawig.so (hand-written HIP for gfx950, C ABI in include/awig.h); PyTorch-
ROCm only allocates HBM and provides the HIP stream.
"""
__version__ = "0.1.0"


def __getattr__(name):
    # lazy imports so that ``import accuratewigner_amd`` needs neither a GPU nor the library
    if name == "WignerMoyalCUDA1D":
        from .wigner_moyal_cuda_1d import WignerMoyalCUDA1D
        return WignerMoyalCUDA1D
    if name == "WignerBlochCUDA1D":
        from .wigner_bloch_cuda_1d import WignerBlochCUDA1D
        return WignerBlochCUDA1D
    if name == "RhoVNeumannCUDA1D":
        from .rho_vneumann_cuda_1d import RhoVNeumannCUDA1D
        return RhoVNeumannCUDA1D
    if name == "RhoBlochCUDA1D":
        from .rho_bloch_cuda_1d import RhoBlochCUDA1D
        return RhoBlochCUDA1D
    if name == "SchrodingerWignerCUDA1D":
        from .schrodinger_wigner_cuda_1d import SchrodingerWignerCUDA1D
        return SchrodingerWignerCUDA1D
    raise AttributeError(name)
