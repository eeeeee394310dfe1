"""GPU: a raw k-space HDF5 file in the reference's layout feeds the adjoint
NUFFT, as in the reference's raw-data reconstruction (``test_interface.cpp:45-150``:
coordinates scaled by 3.141592/160, input = weights * coil data, type 1 per
coil).  Checked against the FP64 direct NDFT at the file's coordinates."""
import numpy as np
import pytest
import torch

from oracle.ndft import ndft_type1, rel_l2
from tests.gpu_util import cuda, host, report

pytestmark = pytest.mark.gpu


def test_raw_kspace_file_feeds_the_adjoint_nufft(tmp_path):
    from hastycucompute_amd import fft as hf
    from hastycucompute_amd import hdf5, synth
    rng = np.random.default_rng(41)
    shape, C = (16, 16, 16), 3
    k = synth.kooshball(96, 32)                                   # radians
    M = len(k[0])
    units = 160.0 / 3.141592                                       # the file's coordinate units
    ts = {f"/Kdata/K{a}_E0": torch.from_numpy((k[i] * units).astype(np.float32).reshape(96, 32))
          for i, a in enumerate("XYZ")}
    ts["/Kdata/KW_E0"] = torch.from_numpy(rng.uniform(0.5, 1.5, M).astype(np.float32))
    for c in range(C):
        ts[f"/Kdata/KData_E0_C{c}"] = torch.from_numpy(
            (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64))
    path = str(tmp_path / "raw.h5")
    hdf5.export_tensors(ts, path)
    (enc,) = hdf5.import_raw_kspace(path)
    assert enc.kdata.shape == (C, M)
    coords = [x.numpy() for x in enc.coords]
    data = (enc.weights[None] * enc.kdata).contiguous()
    plan = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=C, tol=1e-5, sign=1), hf.nufft_type.TYPE_1)
    plan.setpts([cuda(x) for x in coords])
    out = torch.empty((C,) + shape, dtype=torch.complex64, device="cuda")
    plan.execute(data.cuda(), out)
    got = host(out)
    for c in range(C):
        ref = ndft_type1(data[c].numpy(), coords, shape)
        report(f"raw k-space file -> type 1, coil {c}", rel_l2(got[c], ref), 2e-5)
