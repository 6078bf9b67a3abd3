"""The resident p' SpMV on device buffers (cfdhip_pressure_spmv_device).

The 2D stencil kernel must give the same bits as the host-buffer entry point
(itself pinned bitwise to the reference's SparseMatrix::multiply on the golden
systems in test_gpu_parity.py) and as the oracle's CSR SpMV, on a ragged
grid (not a multiple of the 64 x 16 block) with the pinned cell and on a
channel without one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,nx,ny", [("cavity", 130, 70), ("channel", 67, 33)])
def test_device_spmv_bitwise(kind, nx, ny):
    import torch
    import cfdapp_amd as cfd
    from cfdapp_amd import learned
    d = learned.make_case(kind, nx, 500.0, 1.0, ny=ny)
    try:
        rng = np.random.default_rng(nx)
        d.upload("RESPONSE", 0.5 + rng.random(d.n_cells))
        d.upload("FLUX_STAR", rng.uniform(-1, 1, d.n_faces))
        d.assemble_pressure_correction_resident(0)
        x = rng.uniform(-1, 1, d.n_cells)
        y_host = d.pressure_spmv(x)
        rp, col, val = d.pressure_csr()
        y_csr = np.array([np.dot(val[rp[r]:rp[r + 1]], x[col[rp[r]:rp[r + 1]]])
                          for r in range(d.n_cells)])
        xd = torch.from_numpy(x).cuda()
        yd = torch.empty_like(xd)
        torch.cuda.synchronize()
        d.set_timing(True)
        d._ctx.call("cfdhip_pressure_spmv_device", xd.data_ptr(), yd.data_ptr(), 3)
        d._ctx.call("cfdhip_synchronize")
        kt = d.kernel_timing()
        d.set_timing(False)
        y_dev = yd.cpu().numpy()
        assert np.array_equal(y_dev.view(np.uint64), y_host.view(np.uint64))
        # ascending column order from 0.0, as the CSR row: the same bits
        assert np.allclose(y_csr, y_host, rtol=0, atol=1e-14 * np.abs(y_host).max())
        assert kt["launches"] == 3 and kt["mean_us"] > 0
    finally:
        d.close()
