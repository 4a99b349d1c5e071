"""MI355X-native joint motion/image reconstruction (gfx950 HIP kernels).

Python mirror of the reference's hot-path API, over libmeamd.so (C ABI in
include/motion_estimation_amd.h):

* transform: RigidTransformCudaOptimzied, RigidTransformDerivativeCuda
* utils: compute_transform_grids_voxel, generate_motion_parameters
* image_estimation: estimate_image_cg (CG-SENSE)
* transform_estimation: estimate_transform (per-shot Levenberg-Marquardt)
* joint_recon: JointRecon
* pyramid: pyramid_reconstruction, aggregate_sampling_masks, aggregate_transforms
* sense: SenseRecon (the uncorrected SENSE baseline)
* calibration: noise_whiten, coil_compression; espirit: rx_espirit_3d
* io: shot masks from scanner counters, the real-data calibration chain,
  NIfTI-1 and npz outputs; metrics: error volume, gradient entropy
* autograd: DataConsistencyLoss (gradients w.r.t. image and motion)
* batched: BatchedJointRecon (many volumes, HIP graph replay, line-pass tuning)
* simulation: DISORDER sampling, corrupted k-space, the reference's seeded
  motion / noise / sampling generators, run_experiment (the synthetic
  experiment), synthetic benchmark inputs

The package directory is named ``motion-estimation_amd``; load it with
``__graft_entry__.load_package()`` (registers it as ``motion_estimation_amd``).
"""
from . import _lib  # noqa: F401
from .utils import compute_transform_grids_voxel, generate_motion_parameters  # noqa: F401
from .transform import RigidTransformCudaOptimzied, RigidTransformDerivativeCuda  # noqa: F401
from . import encoding, simulation  # noqa: F401
from .image_estimation import estimate_image_cg  # noqa: F401
from .transform_estimation import estimate_transform  # noqa: F401
from .joint_recon import JointRecon  # noqa: F401
from .pyramid import pyramid_reconstruction, aggregate_sampling_masks, aggregate_transforms  # noqa: F401
from .autograd import DataConsistencyLoss, data_consistency_loss  # noqa: F401
from .sense import SenseRecon, sense_recon  # noqa: F401
from .calibration import coil_compression, noise_whiten  # noqa: F401
from .espirit import rx_espirit_3d  # noqa: F401
from .batched import BatchedJointRecon  # noqa: F401
from . import io, metrics  # noqa: F401
