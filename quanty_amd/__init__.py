"""MI355X engine for Quanty's second-quantization hot path (H*psi, Lanczos,
block Lanczos, lowest eigenstates, continued-fraction spectra)."""
from ._lib import Context, QSError, default_context, load  # noqa: F401
