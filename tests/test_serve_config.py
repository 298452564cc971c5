"""Request validation of the server (libspimage_amd/serve.py) before any
device work: shapes whose axes have no Stockham plan (which batches do not
take), iteration counts off the quantum, unknown algorithms."""
import numpy as np
import pytest


def _req(shape, **kw):
    from libspimage_amd.serve import Request
    z = np.zeros(shape, np.float32)
    args = dict(rid=7, amplitudes=z, mask=z, support=z, algorithm="hio", iterations=100, seed=1)
    args.update(kw)
    return Request(**args)


@pytest.mark.parametrize("shape,kw,msg", [
    ((47, 64), {}, "Stockham plan"),
    ((64, 64, 11), {}, "Stockham plan"),
    ((64, 64), {"iterations": 70}, "multiple of 50"),
    ((64, 64), {"algorithm": "dm"}, "unknown algorithm"),
])
def test_requests_refused_before_serving(shape, kw, msg):
    from libspimage_amd.serve import Server
    srv = Server()
    with pytest.raises(ValueError, match=msg):
        srv.run([_req(shape, **kw)])
    assert srv.batches_created == 0
