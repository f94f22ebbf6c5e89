"""The library's kernels on the CPU, under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cpu_emu/): the unmodified .hip sources
are compiled with g++ over an emulated HIP runtime (workgroup threads as
fibers, wave intrinsics as rendezvous of the wave's lanes) and run through
the C ABI against the oracle -- builds of every builder and layout, the
three traversals with FCP and kNN lists, Morton-sorted and allocation-free
launches, forests and spatial trees.  A second binary builds with the
finish-capacity thresholds at 1, so that these small inputs run the
largest LDS capacities.  An out-of-bounds access or undefined behaviour in
any kernel aborts with the sanitizer's report; a result that differs from
the oracle fails the case.  The builds also run with skewed wave
schedules (EMU_SKEW): one wave at a time runs on to the next workgroup
barrier, so that races between waves that drift apart (wave-local phases)
change the result."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emu")


def test_kernels_under_sanitizers_match_oracle():
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    subprocess.run(["make", "-s", "-j", jobs, "-C", EMU], check=True, timeout=1200)
    for name in ("emu_test", "emu_test_bigcap"):
        r = subprocess.run([os.path.join(EMU, "_build", name)], capture_output=True, text=True,
                           timeout=1200)
        print(r.stdout[-4000:])
        assert r.returncode == 0 and "ALL OK" in r.stdout, \
            "%s failed:\n%s\n%s" % (name, r.stdout[-3000:], r.stderr[-6000:])


def test_builds_under_skewed_wave_schedules():
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    subprocess.run(["make", "-s", "-j", jobs, "-C", EMU], check=True, timeout=1200)
    procs = []
    for name, skew in (("emu_test", "1"), ("emu_test_bigcap", "2")):
        env = dict(os.environ, EMU_SKEW=skew)
        procs.append((name, skew, subprocess.Popen([os.path.join(EMU, "_build", name), "build"],
                                                   stdout=subprocess.PIPE,
                                                   stderr=subprocess.PIPE, text=True, env=env)))
    for name, skew, p in procs:
        out, err = p.communicate(timeout=1200)
        assert p.returncode == 0 and "ALL OK" in out, \
            "%s EMU_SKEW=%s failed:\n%s\n%s" % (name, skew, out[-3000:], err[-6000:])
