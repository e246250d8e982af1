"""distCUDA2 of the reference's compiled ``simple_knn._C`` (ext.cpp:15-17),
served by the HIP library."""
from partgs_amd.knn import distCUDA2  # noqa: F401
