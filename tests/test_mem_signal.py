"""bh_mem_signal (include/bh_mem_signal.h): the SIGSEGV dispatcher that replaces the reference's
libsigsegv-based one (core/bh_mem_signal.cpp:121-196).  A small C program linked against the built
libbh.so checks attach/detach/exist, faults on protected and read-only pages, and the chaining of
faults outside every segment to the handler installed before (or after) it.  CPU only."""
import os
import signal
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "bohrium_amd", "lib")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("mem_signal") / "mem_signal_check")
    subprocess.run(["gcc", "-O1", "-std=gnu99", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "csrc", "mem_signal_check.c"), "-o", exe, "-L" + LIB, "-lbh",
                    "-Wl,-rpath," + LIB], check=True)
    return exe


def _run(exe, mode):
    return subprocess.run([exe, mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)


def test_segments_and_faults(checker):
    r = _run(checker, "basic")
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout


def test_fault_outside_segments_is_fatal(checker):
    r = _run(checker, "foreign")
    assert r.returncode == -signal.SIGSEGV, (r.returncode, r.stdout)


def test_fault_outside_segments_reaches_earlier_handler(checker):
    r = _run(checker, "chained")
    assert r.returncode == 42, (r.returncode, r.stdout)


def test_later_handler_is_put_behind(checker):
    r = _run(checker, "reinstall")
    assert r.returncode == 43, (r.returncode, r.stdout)
