"""The drop-in under the reference's own names, on the CPU: ``import gpuRIR`` exposes the
reference module's API (gpuRIR/__init__.py:11-246) and ``gpuRIR_bind`` the compiled class's
interface (src/python_bind.cpp:130-143), with the constructor's and toggles' messages
(src/gpuRIR_cuda.cu:1025-1073).  No GPU work: the library is only loaded."""
import inspect

import gpuRIR
import gpuRIR_bind
import gpurir_amd


def test_gpuRIR_exports_the_reference_api():
    for name in ["polar_patterns", "beta_SabineEstimation", "att2t_SabineEstimator", "t2n", "simulateRIR",
                 "simulateTrajectory", "activateMixedPrecision", "activateLUT", "activate_mixed_precision",
                 "activate_lut"]:
        assert getattr(gpuRIR, name) is getattr(gpurir_amd, name)
    ns = {}
    exec("from gpuRIR import *", ns)  # the reference's __all__ names all resolve here
    assert set(gpuRIR.__all__) <= set(ns)
    assert gpuRIR.polar_patterns == {"omni": 0, "homni": 1, "card": 2, "hypcard": 3, "subcard": 4, "bidir": 5}
    sig = inspect.signature(gpuRIR.simulateRIR)
    assert list(sig.parameters) == ["room_sz", "beta", "pos_src", "pos_rcv", "nb_img", "Tmax", "fs", "Tdiff",
                                    "spkr_pattern", "mic_pattern", "orV_src", "orV_rcv", "c"]
    assert list(inspect.signature(gpuRIR.simulateTrajectory).parameters) == ["source_signal", "RIRs", "timestamps",
                                                                            "fs"]
    import gpuRIR.torch_ops as to
    assert to.simulate_batch_diff is gpurir_amd.torch_ops.simulate_batch_diff


def test_bind_class_interface():
    sim = gpuRIR_bind.gpuRIR_bind
    for m in ["simulateRIR_bind", "gpu_conv", "activate_mixed_precision_bind", "activate_lut_bind"]:
        assert callable(getattr(sim, m))
    params = list(inspect.signature(sim.simulateRIR_bind).parameters)[1:]
    assert params == ["room_sz", "beta", "pos_src", "pos_rcv", "orV_src", "orV_rcv", "spkr_pattern", "mic_pattern",
                      "nb_img", "Tdiff", "Tmax", "Fs", "c"]
    assert inspect.signature(sim.simulateRIR_bind).parameters["c"].default == 343.0
    assert list(inspect.signature(sim.gpu_conv).parameters)[1:] == ["source_segments", "RIR"]


def test_bind_flags_and_messages(capsys):
    s = gpuRIR_bind.gpuRIR_bind()
    assert (s.mixed_precision, s.lookup_table) == (False, True)  # python_bind.cpp:15 defaults
    assert capsys.readouterr().out == ""
    assert s.activate_mixed_precision_bind(True) is True
    assert s.lookup_table is False
    assert capsys.readouterr().out == ("The mixed precision implementation is not compatible with the lookup table.\n"
                                       "Disabling the lookup table.\n")
    assert s.activate_lut_bind(True) is True and s.mixed_precision is False
    assert capsys.readouterr().out == ("The lookup table is not compatible with the mixed precision implementation.\n"
                                       "Disabling the mixed precision implementation.\n")
    assert s.activate_lut_bind(False) is False
    # the constructor activates mixed precision, then the table (src/gpuRIR_cuda.cu:1031-1033)
    t = gpuRIR_bind.gpuRIR_bind(True)
    assert (t.mixed_precision, t.lookup_table) == (False, True)
    assert "Disabling the mixed precision implementation." in capsys.readouterr().out
    # instances keep their own flags
    assert (s.mixed_precision, s.lookup_table) == (False, False)
