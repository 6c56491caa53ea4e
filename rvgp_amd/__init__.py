"""rvgp_amd -- the RVGP connection-Laplacian GP path on AMD MI355X.

Drop-in for the reference package's public API (RVGP/__init__.py:1-3):

    import rvgp_amd as RVGP
    from rvgp_amd.geometry import furthest_point_sampling
    d = RVGP.create_data_object(X, n_eigenpairs=50)
    d.random_vector_field(seed=1); d.smooth_vector_field(t=100)
    gp = RVGP.fit(d, train_ind=train_ind, noise_variance=0.001)
    mean, var = gp.transform(d, test_ind)

All numerical work runs in the HIP kernels of librvgp.so (include/rvgp.h);
importing the package does not touch the GPU.
"""
__version__ = "0.1.0"


def __getattr__(name):
    # Lazy so that `import rvgp_amd` works on a machine without a GPU.
    if name == "create_data_object":
        from .dataclass import data
        return data
    if name == "fit":
        from .main import train_gp
        return train_gp
    if name == "fit_batched":
        from .batch import fit_batched
        return fit_batched
    if name == "kernels":
        from . import kernels
        return kernels
    raise AttributeError(name)
