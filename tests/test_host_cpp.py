"""C++ host side: the RadiationModel-style interface (tests/cpp/test_radiation_model.cpp) and the
helios::RayTracingBackend adapter (helios_amd/host/HipRayTracingBackend.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_radiation_model")
REF = "/root/reference"


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return BIN


def test_cpp_host_cases_cpu():
    r = subprocess.run([_binary(), "--cpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "plugins/radiation/include/RayTracingBackend.h")),
                    reason="Helios source tree not present")
def test_adapter_compiles_against_reference_interface():
    """The adapter overrides every pure virtual of helios::RayTracingBackend with the reference's exact
    signatures: it must compile (and instantiate) against the reference's own headers."""
    src = os.path.join(ROOT, "helios_amd", "host", "HipRayTracingBackend.cpp")
    probe = ("#include \"HipRayTracingBackend.h\"\n"
             "int main(){ helios::HipRayTracingBackend b(0); helios::RayTracingBackend *p = &b; (void)p; return 0; }\n")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "helios_amd", "host"),
           "-I", os.path.join(REF, "core/include"), "-I", os.path.join(REF, "core/lib/pugixml"),
           "-I", os.path.join(REF, "plugins/radiation/include")]
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror=overloaded-virtual", *inc, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", *inc, "-x", "c++", "-"], input=probe,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "plugins/radiation/src/RayTracingTypes.cpp")),
                    reason="Helios source tree not present")
def test_adapter_runs_against_recording_abi(tmp_path):
    """The adapter runs through the reference's RayTracingBackend interface on the CPU, over a recording test
    double of the C ABI (tests/cpp/adapter_fake_abi.cpp): geometry/mask/material/source conversion, the band
    table with retired bands, launch parameters, results and errors (tests/cpp/test_adapter.cpp).  The binary
    is built in a temporary directory from the reference's RayTracingTypes.cpp where it lies and never leaves
    this machine."""
    exe = str(tmp_path / "test_adapter")
    cpp = os.path.join(ROOT, "tests", "cpp")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "helios_amd", "host"), "-I", cpp,
           "-I", os.path.join(REF, "core/include"), "-I", os.path.join(REF, "core/lib/pugixml"),
           "-I", os.path.join(REF, "plugins/radiation/include")]
    srcs = [os.path.join(ROOT, "helios_amd", "host", "HipRayTracingBackend.cpp"), os.path.join(cpp, "adapter_fake_abi.cpp"),
            os.path.join(cpp, "test_adapter.cpp"), os.path.join(REF, "plugins/radiation/src/RayTracingTypes.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-O1", *inc, *srcs, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "all checks passed" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_host_cases_gpu():
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
