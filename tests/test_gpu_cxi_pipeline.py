"""Patterns in from a CXI file, reconstructions and their PRTF out to one:
the reading and writing change nothing.  Each event read back equals the
array written, the Reconstructor run on it equals the run on the in-memory
array bit for bit, and the written results read back equal."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_cxi_in_reconstruct_prtf_cxi_out(sp, tmp_path):
    from libspimage_amd import cxi, workloads
    from libspimage_amd.prtf import prtf
    from libspimage_amd.reconstructor import Reconstructor
    try:
        cxi._lib()
    except OSError as e:
        pytest.skip(f"libhdf5 unavailable: {e}")
    shape = (64, 64)
    pats = [workloads.make_pattern(shape, n_gauss=6, radius=(10, 10), photons=1e6, beamstop=2, seed=s) for s in (5, 6)]
    intens = np.array([np.fft.fftshift(p.amplitudes.astype(np.float64) ** 2) for p in pats])
    masks = np.array([np.fft.fftshift(p.mask.astype(bool)) for p in pats])
    src = tmp_path / "patterns.cxi"
    w = cxi.CXIWriter(str(src), 2)
    w.write_stack({"entry_1": {"data_1": {"data": intens, "mask": masks}}})
    w.close()

    def configure(I, M):
        R = Reconstructor()
        R.set_intensities(I)
        R.set_mask(M)
        R.set_number_of_iterations(60)
        R.set_number_of_outputs(3)
        R.set_initial_support(radius=12)
        R.set_support_algorithm("area", update_period=10, blur_init=3.0, blur_final=1.0, area_init=0.12,
                                area_final=0.08)
        R.append_phasing_algorithm("raar", number_of_iterations=45, beta_init=0.9, beta_final=0.8)
        R.append_phasing_algorithm("er", number_of_iterations=15)
        return R

    r = cxi.CXIReader(str(src), {"I": "/entry_1/data_1/data", "M": "/entry_1/data_1/mask"},
                      def_stack_ds="/entry_1/data_1/data")
    out_path = tmp_path / "results.cxi"
    ow = cxi.CXIWriter(str(out_path), 8)
    k = 0
    while (ev := r.get_next()) is not None:
        assert np.array_equal(ev["I"], intens[ev["i"]])
        assert np.array_equal(ev["M"].astype(bool), masks[ev["i"]])
        res = configure(ev["I"], ev["M"].astype(bool)).reconstruct_loop(4, seed=100)
        ref = configure(intens[ev["i"]], masks[ev["i"]]).reconstruct_loop(4, seed=100)
        assert np.array_equal(res["real_space_final"], ref["real_space_final"])
        p = prtf(res["real_space_final"], res["support_final"], full_out=True)
        for j in range(4):
            ow.write_slice({"entry_1": {"image_1": {"data": res["real_space_final"][j].astype(np.complex64),
                                                    "mask": res["support_final"][j].astype(np.int8)}}}, i=k)
            k += 1
        ow.write_slice({f"prtf_{ev['i']}": {"prtf": p["prtf"].astype(np.float32),
                                            "super_image": p["super_image"].astype(np.complex64)}})
    ow.close()
    with cxi.H5File(str(out_path)) as f:
        imgs = f.read("/entry_1/image_1/data")
        assert imgs.shape == (8,) + shape and imgs.dtype == np.complex64
        assert np.array_equal(imgs[4:], configure(intens[1], masks[1]).reconstruct_loop(4, seed=100)
                              ["real_space_final"].astype(np.complex64))
        p1 = f.read("/prtf_1/prtf")
        assert p1.shape == shape and np.all((p1 >= 0) & (p1 <= 1 + 1e-6))
