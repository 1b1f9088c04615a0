"""neb_amd -- MI355X (gfx950) backend for the neural eigenfluid network hot path.

Analytic second-order spatial jets of a shared tanh/sin trunk, the lifted
velocity / vorticity / velocity-gradient fields, quadrature-weighted FP64
Grams Phi^T W Phi, the spectral objective and the handwritten parameter
reverse pass, as hand-written HIP kernels behind the C ABI of
include/neb_hip.h (libnebhip.so).
"""
from .network import (EigenfluidNet, Fields, NetConfig, Tape, backward_fields, backward_jets, channels, cross,
                      envelope_jets, gram,
                      output_fields, output_jets, trunk_forward)
from .modes import ModeConfig, ModeNet
from .spectral import (JetGram, PeelSolver, SpectralSolver, basis_stream, chunked_step, coefficient_peel, compute,
                       streamfunction_peel_step, velocity_peel_step,
                       fullmass, generalized_eigh, project, project_streamed, spectral_objective)

__all__ = ["EigenfluidNet", "Fields", "NetConfig", "Tape", "backward_fields", "backward_jets", "envelope_jets", "channels", "cross", "gram",
           "output_fields", "output_jets", "trunk_forward", "JetGram", "PeelSolver", "SpectralSolver", "coefficient_peel", "basis_stream", "chunked_step", "compute", "fullmass", "generalized_eigh", "project", "project_streamed",
           "spectral_objective", "streamfunction_peel_step", "velocity_peel_step", "ModeConfig", "ModeNet"]
