"""treecorr_amd: TreeCorr's ball-tree correlation engine on AMD MI355X (gfx950).

Drop-in for the hot path of TreeCorr 5.1.3: Catalog/Field and the NN/KK/GG
2-point and NNN/GGG multipole 3-point correlations, computed by hand-written HIP
kernels behind a C ABI (include/treecorr_amd.h).
"""
from .catalog import (Catalog, calculateVarG, calculateVarK, calculateVarQ, calculateVarT,
                      calculateVarV, calculateVarZ)
from .field import Field, NField, KField, GField
from .corr2 import (Corr2, NNCorrelation, KKCorrelation, GGCorrelation, NKCorrelation,
                    NGCorrelation, KGCorrelation, ZZCorrelation, NZCorrelation, KZCorrelation,
                    VVCorrelation, TTCorrelation, QQCorrelation, NVCorrelation,
                    NTCorrelation, NQCorrelation, KVCorrelation, KTCorrelation,
                    KQCorrelation)
from .util import set_omp_threads, get_omp_threads

from .corr3 import Corr3, NNNCorrelation, GGGCorrelation

__version__ = '0.1.0'
