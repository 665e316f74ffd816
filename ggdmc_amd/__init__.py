"""ggdmc_amd: MI355X-native DE-MCMC engine for LBA models (drop-in for ggdmc's
src/de.cpp sampler path). Compute runs in libggdmc_hip.so (HIP, gfx950)."""
from ._lib import (FP64_BATCHED, FP64_PLAIN, MIXED_BATCHED, KIND_HIER, KIND_HYPER,  # noqa: F401
                   KIND_SUBJECT, GgdmcError, device_count)
from .runs import (JobSpec, LBAModel, Posterior, PriorSpec, Session, SubjectData,  # noqa: F401
                   loglike, make_config, run, run_batch, run_hyper, run_subject)
