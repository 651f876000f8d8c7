"""gkmsvm_amd: an MI355X (gfx950) engine for gapped k-mer SVM kernels.

Drop-in for gkmsvm-lite's scoring, variant scoring, GkmExplain, ISM, the
Gram build behind train_gkmsvm, the exact l-mer tables and to_deltasvm; the
work runs in hand-written HIP kernels
(libgkmhip.so, C-ABI in include/gkmhip.h) called through ctypes.
"""

from .codec import (
    codes_to_one_hot,
    encode_batch,
    encode_codes,
    one_hot_decode,
    one_hot_encode,
    reverse_complement,
    to_codes,
    validate,
)
from .deltasvm import DeltaSVM
from .explain import gkmexplain, mutation_impact
from .gram import compute_gram
from .importers import (
    load_classic_model,
    load_deltasvm_model,
    load_lsgkm_model,
    load_r_gkmsvm_model,
    read_fasta,
)
from .ism import ism
from .kernels import (
    KERNEL_ALIASES,
    KERNEL_BUILDERS,
    CenterWeightedGkmKernel,
    CenterWeightedRbfGkmKernel,
    DirectGkmKernel,
    EstTruncGkmKernel,
    RbfGkmKernel,
    RbfKernel,
    resolve_kernel_type,
)
from .serialization import (
    load_deltasvm_npz,
    load_model,
    load_npz,
    save_deltasvm_npz,
    save_lsgkm,
    save_npz,
)
from .svm import GkmSVM, to_cpu
from .train import train_gkmsvm, train_gkmsvr

__all__ = [
    "GkmSVM", "gkmexplain", "mutation_impact", "ism", "compute_gram", "train_gkmsvm",
    "train_gkmsvr",
    "load_lsgkm_model", "read_fasta", "one_hot_encode", "encode_codes", "to_codes",
    "codes_to_one_hot", "to_cpu", "resolve_kernel_type", "KERNEL_BUILDERS", "KERNEL_ALIASES",
    "DirectGkmKernel", "EstTruncGkmKernel", "CenterWeightedGkmKernel", "RbfKernel",
    "RbfGkmKernel", "CenterWeightedRbfGkmKernel", "DeltaSVM", "load_classic_model",
    "load_r_gkmsvm_model", "load_model", "load_npz", "save_npz", "save_lsgkm",
    "load_deltasvm_npz", "save_deltasvm_npz", "load_deltasvm_model", "encode_batch",
    "one_hot_decode", "reverse_complement", "validate",
]
