"""The reference's import paths resolve to the build (CPU: imports only, no solve)."""


def test_import_sigpde_torch_as_sig():
    import sigpde.torch as sig

    import sigpde_amd.torch as amd
    assert sig is amd
    assert sig.SigPDE is amd.SigPDE and sig.RobustSigPDE is amd.RobustSigPDE
    k = sig.kernels.LinearKernel()
    assert k.scale == 1.0
    assert sig.kernels.RBFKernel().sigma == 1.0 and sig.kernels.L1RBFKernel(2.0).sigma == 2.0
    assert sorted(sig.__all__) == ["RobustSigPDE", "SigPDE", "kernels", "normalizers"]


def test_reference_submodule_paths():
    from sigpde.torch.kernel import RobustSigPDE, SigPDE, log_normalizer
    from sigpde.torch.static_kernels import L1RBFKernel, LinearKernel, RBFKernel

    import sigpde
    import sigpde_amd.torch as amd
    assert SigPDE is amd.SigPDE and RobustSigPDE is amd.RobustSigPDE
    assert LinearKernel is amd.kernels.LinearKernel
    assert RBFKernel is amd.kernels.RBFKernel and L1RBFKernel is amd.kernels.L1RBFKernel
    assert callable(log_normalizer)
    assert sigpde.torch is amd
