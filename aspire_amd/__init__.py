"""aspire_amd -- MI355X-native common-lines engine (drop-in for ASPIRE's
``cryo_clmatrix``).  The compute path is libclm.so (HIP, gfx950)."""
from .clmatrix import (assemble, autotune, clmatrix_grouped, clmatrix_records, cryo_clmatrix,  # noqa: F401
                       cryo_clmatrix_subset, partner_subset)
from .pft import cryo_pft  # noqa: F401
from .shifts import cryo_estimate_shifts  # noqa: F401

__all__ = ["cryo_clmatrix", "clmatrix_records", "clmatrix_grouped", "autotune", "assemble", "cryo_pft",
           "cryo_clmatrix_subset", "partner_subset", "cryo_estimate_shifts"]
