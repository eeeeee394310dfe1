"""GPU: a device tensor in the generic_value wire format carries device_type 1
and its device index (generic_value.cppm:459-466), and is read back onto that
device; a SENSE input set (coordinates, k-space, maps) round-trips exactly."""
import pytest
import torch

from hastycucompute_amd import generic_value as gv


@pytest.mark.gpu
def test_device_tensors_round_trip():
    g = torch.Generator().manual_seed(4)
    v = {"coords": [torch.rand(64, generator=g).cuda() for _ in range(3)],
         "kdata": torch.randn(2, 64, generator=g, dtype=torch.complex64).cuda(),
         "maps": torch.randn(2, 4, 4, 4, generator=g, dtype=torch.complex64)}
    b = gv.serialize(v)
    kpos = b.index(b"kdata") + 5
    assert b[kpos + 8] == 1 and b[kpos + 8 + 3] == 0          # device_type cuda, index 0
    back = gv.deserialize(b)
    assert back["kdata"].device == torch.device("cuda:0") and back["maps"].device.type == "cpu"
    assert torch.equal(back["kdata"].cpu(), v["kdata"].cpu())
    assert all(torch.equal(a.cpu(), c.cpu()) for a, c in zip(back["coords"], v["coords"]))
    assert gv.metadata_string(v["kdata"]) == "Tensor[cuda:0,Complex64,shape=(2,64)]"
    host = gv.deserialize(b, device="cpu")
    assert host["kdata"].device.type == "cpu"
