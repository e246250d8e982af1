"""Import name of the reference's simple-knn package, served by partgs_amd
(``from simple_knn._C import distCUDA2``, scene/gaussian_model.py:8)."""
