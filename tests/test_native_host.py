"""The host-side native code under sanitizers, and the JNI shim through a
fake JNIEnv.

CPU: tests/native/host_selftest.cpp (validation, relabelling, the threaded
batch preparation and decoding, certificate batches, the threaded class
generator) built with AddressSanitizer + UndefinedBehaviorSanitizer and with
ThreadSanitizer; tests/native/oracle_selftest.c (the oracle, including its
threaded batch runner) the same way; the JNI shim with the stub C ABI under
ASan + UBSan.  GPU: the JNI shim linked against libtd16.so (built by
__graft_entry__.build() into tests/native/_build)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
CSRC = os.path.join(ROOT, "td16_amd", "csrc")
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
TSAN = ["-fsanitize=thread"]


def build_and_run(tmp_path, compiler, name, sources, flags, expect, env=None):
    exe = str(tmp_path / name)
    cmd = [compiler, "-O1", "-g", "-pthread", "-I" + ROOT] + flags + ["-o", exe] + sources
    if compiler == "g++":
        cmd.insert(1, "-std=c++17")
    subprocess.run(cmd, check=True, capture_output=True)
    run_env = dict(os.environ)
    run_env.update(env or {})
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=run_env)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert expect in out.stdout
    return out


HOST = [os.path.join(NATIVE, "host_selftest.cpp")] + [
    os.path.join(CSRC, f) for f in ("td16_host.cpp", "td16_prepare.cpp", "td16_bracelet.cpp")]
ORACLE = [os.path.join(NATIVE, "oracle_selftest.c"), os.path.join(ROOT, "oracle", "td16_oracle.c"),
          "-lm"]


def test_host_core_under_address_and_undefined_sanitizers(tmp_path):
    build_and_run(tmp_path, "g++", "host_asan", HOST, ASAN, "host selftest ok")


def test_host_core_under_thread_sanitizer(tmp_path):
    build_and_run(tmp_path, "g++", "host_tsan", HOST, TSAN, "host selftest ok",
                  env={"TSAN_OPTIONS": "halt_on_error=1"})


def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    build_and_run(tmp_path, "gcc", "oracle_asan", ORACLE, ASAN, "oracle selftest ok")


def test_oracle_under_thread_sanitizer(tmp_path):
    build_and_run(tmp_path, "gcc", "oracle_tsan", ORACLE, TSAN, "oracle selftest ok",
                  env={"TSAN_OPTIONS": "halt_on_error=1"})


def test_jni_shim_marshalling_through_a_fake_env(tmp_path):
    """jni_sorter.cpp compiled against the test stand-in for jni.h and a stub
    C ABI: int[len][3] positions, int[0][] without a witness, null plus
    RuntimeException(message) on errors, unsigned bytes, JNI_ABORT releases,
    the served mode's options."""
    srcs = [os.path.join(NATIVE, "jni_driver.cpp"), os.path.join(CSRC, "jni_sorter.cpp"),
            os.path.join(NATIVE, "td16_stub.cpp")]
    flags = ASAN + ["-DTD16_JNI_STUB_ABI", "-I" + os.path.join(NATIVE, "jni_stand_in")]
    build_and_run(tmp_path, "g++", "jni_stub", srcs, flags, "jni driver ok")


@pytest.mark.gpu
def test_jni_shim_on_the_gpu():
    """The JNI shim on the real library: the reference's smoke vector, a
    product of six transpositions on twelve symbols (a witness that replays)
    and an invalid input (RuntimeException), through the fake JNIEnv."""
    exe = os.path.join(NATIVE, "_build", "jni_selftest")
    assert os.path.exists(exe), "build it with __graft_entry__.build()"
    print("case: jni shim smoke + n=12 product + invalid input", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert "jni driver ok" in out.stdout
