"""MI355X-native dictionary indexing for EBSD with ebsdtorch's call surface.

Compute runs in libebsd_amd.so (hand-written gfx950 HIP kernels behind the C
ABI of include/ebsd_amd.h); the Python layer mirrors ebsdtorch's classes and
drivers for the dictionary-indexing path.
"""

from .experiment_pats import ExperimentPatterns
from .fitting import spatial_fft_pc_fitter
from .geometry import EBSDGeometry, bruker_geometry_to_SE3
from .indexing import (DictionaryIndexer, PCADictionaryIndexer, dictionary_index_orientations,
                       pca_dictionary_index_orientations)
from .io import read_master_pattern
from .pca_covmat import OnlineCovMatrix
from .knn import ChunkedKNN
from .master_pattern import MasterPattern, point_group_to_laue, space_group_to_laue
from .orientations import (
    disori_angle_laue,
    get_laue_mult,
    laue_elements,
    ori_in_fz_laue,
    ori_to_fz_laue,
    qu_apply,
    sample_ori_fz_laue,
    sample_ori_fz_laue_angle,
    so3_cubochoric_grid,
    so3_uniform_quat,
)
from .preprocessing import clahe_grayscale, nlpar
from .refining import orientation_grad_refinement, orientation_grid_refinement

__all__ = [
    "ChunkedKNN", "DictionaryIndexer", "EBSDGeometry", "ExperimentPatterns", "MasterPattern", "OnlineCovMatrix",
    "PCADictionaryIndexer", "bruker_geometry_to_SE3", "clahe_grayscale", "dictionary_index_orientations",
    "disori_angle_laue", "get_laue_mult", "laue_elements", "nlpar", "ori_in_fz_laue", "ori_to_fz_laue",
    "orientation_grad_refinement", "orientation_grid_refinement", "pca_dictionary_index_orientations",
    "point_group_to_laue", "qu_apply", "read_master_pattern", "sample_ori_fz_laue", "sample_ori_fz_laue_angle",
    "so3_cubochoric_grid", "so3_uniform_quat", "space_group_to_laue", "spatial_fft_pc_fitter",
]
