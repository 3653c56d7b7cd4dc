"""fast-er for MI355X: the Fast-ER linkage pipeline on gfx950 HIP kernels.

Drop-in for the ``faster`` package of fast-er 0.3.1: the same classes with the
same constructors, ``fit``/``transform`` signatures and outputs. Pair lists
are int64 torch tensors on the current ROCm device.
"""
__version__ = '0.3.1'

from .comparison import Comparison
from .deduplication import Deduplication
from .estimation import Estimation
from .linkage import Linkage
from .evaluation import Evaluation
