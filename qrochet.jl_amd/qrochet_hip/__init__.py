"""qrochet_hip: Qrochet.jl's MPS (`Chain`) update path on MI355X (gfx950).

Python mirror of the reference's Julia API over libqrochet_hip.so.  Julia
`f!` methods are spelled `f_`.
"""
from .chain import (  # noqa: F401
    MPS,
    Chain,
    Context,
    Dense,
    Operator,
    Site,
    State,
    canonize,
    canonize_,
    canonize_site,
    canonize_site_,
    contract_between,
    contract_between_,
    default_context,
    evolve_,
    evolve_layer_,
    expect,
    expect_many,
    from_internal,
    gate,
    isleftcanonical,
    isrightcanonical,
    mixed_canonize,
    mixed_canonize_,
    norm,
    normalize_,
    overlap,
    overlap_noalloc,
    pMPS,
    rand_chain,
    simple_update_,
    site,
    truncate,
    truncate_,
)
from .errors import (  # noqa: F401
    AdmissionError,
    ArgumentError,
    DimensionMismatch,
    MissingSchmidtCoefficientsException,
    QHError,
)
