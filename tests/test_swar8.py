"""The byte-parallel helpers of pargo_amd/csrc/swar8.h against per-byte
scalar statements (every byte value in every byte position), and the scad8
init pass for interior tiles in pargo_amd/csrc/sca_swar.h against the
per-cell statements of the general path on random tiles of every class
kind, both compiled for the host with g++."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def _build_and_run(name):
    src = os.path.join(HERE, "cpp", name + ".cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, name)
        subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-unknown-pragmas", src, "-o", exe], check=True)
        out = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
        assert out.returncode == 0, out.stdout
        assert out.stdout.strip() == "ok"


def test_swar8_helpers_match_bytewise_statements():
    _build_and_run("swar8_check")


def test_sca_init_words_match_per_cell_statements():
    _build_and_run("sca_swar_check")
