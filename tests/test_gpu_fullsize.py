"""The projector pair at the full benchmark size (256^3 voxels x 360 angles, one
emission line, 24 GB of attenuation), checked through properties that hold at
any size: the adjoint identity <F C, r> = <C, B^T r> (FP64 sums of the FP32
outputs), linearity in C and in the ratio, and bit-for-bit repeatability.
The oracle is too slow at this size; the same kernels are checked against it
at smaller sizes in test_gpu_projector.py."""
import pytest
import torch

pytestmark = pytest.mark.gpu

N, A = 256, 360


@pytest.fixture(scope="module")
def data():
    g = torch.Generator(device="cuda")
    g.manual_seed(20260002)
    atten = torch.empty((A, N, N, N), device="cuda").uniform_(0.3, 1.0, generator=g)
    C = torch.rand((1, N, N, N), device="cuda", generator=g)
    C2 = torch.rand((1, N, N, N), device="cuda", generator=g)
    em = torch.full((A, 1), 1e-4, device="cuda")
    theta = torch.deg2rad(torch.arange(A, device="cuda", dtype=torch.float64) * (180.0 / A)).float()
    r = torch.rand((A, N, N), device="cuda", generator=g) + 0.5
    yield atten, C, C2, em, theta, r
    del atten
    torch.cuda.empty_cache()


def test_adjoint_identity_full_size(data):
    from fl_correction_amd.cuda_lib import atten_cuda
    atten, C, _, em, theta, r = data
    print(f"full size: ({A}, {N}, {N}, {N}), adjoint identity", flush=True)
    Pf = atten_cuda.forward_emission_batch(atten, C, em, theta)
    Bt = atten_cuda.backward_emission_batch(atten, r, em, theta)
    lhs = float((Pf.double() * r.double()).sum())
    rhs = float((C.double() * Bt.double()).sum())
    print(f"  <FC, r> = {lhs:.10e}, <C, B^T r> = {rhs:.10e}, rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_linearity_and_repeatability_full_size(data):
    from fl_correction_amd.cuda_lib import atten_cuda
    atten, C, C2, em, theta, r = data
    print(f"full size: ({A}, {N}, {N}, {N}), linearity", flush=True)
    P1 = atten_cuda.forward_emission_batch(atten, C, em, theta)
    P2 = atten_cuda.forward_emission_batch(atten, C2, em, theta)
    P12 = atten_cuda.forward_emission_batch(atten, (C + 2.0 * C2).contiguous(), em, theta)
    e = float(torch.linalg.vector_norm((P1 + 2.0 * P2 - P12).double()) / torch.linalg.vector_norm(P12.double()))
    print(f"  forward linearity rel-L2 {e:.2e}")
    assert e <= 1e-6
    B1 = atten_cuda.backward_emission_batch(atten, r, em, theta)
    B2 = atten_cuda.backward_emission_batch(atten, (3.0 * r).contiguous(), em, theta)
    e = float(torch.linalg.vector_norm((3.0 * B1 - B2).double()) / torch.linalg.vector_norm(B2.double()))
    print(f"  back-projection homogeneity rel-L2 {e:.2e}")
    assert e <= 1e-6
    assert torch.equal(P1, atten_cuda.forward_emission_batch(atten, C, em, theta))
    assert torch.equal(B1, atten_cuda.backward_emission_batch(atten, r, em, theta))
