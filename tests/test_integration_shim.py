"""The reference-side binding shown in INTEGRATION.md (section 2) compiles against the
reference's own headers and links, with no undefined symbol, against the reference objects
(mc_main.cc and everything it links, built from the upstream sources by oracle/ref.mk) and
the product library.  Needs the reference tree, so it runs in the build container only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not on this machine")
def test_documented_shim_links_against_the_reference():
    lib = os.path.join(ROOT, "moribs-pimc_amd", "libmoribs_pimc.so")
    if not os.path.exists(lib):
        pytest.skip("product library not built")
    r = subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref.mk"), f"REF={REF}", "shim"],
                       cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    so = os.path.join(ROOT, "oracle", "_ref", "libgpu_shim.so")
    syms = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    assert "moribs_shim_entries" in syms
    # the shim's calls into the product resolve to the library's exports
    und = subprocess.run(["nm", "-D", "--undefined-only", so], capture_output=True, text=True).stdout
    for name in ("pimc_create", "pimc_set_replica", "pimc_commit", "pimc_set_state", "pimc_get_state",
                 "pimc_get_counters", "pimc_add_pot1d", "pimc_add_pot2d", "pimc_add_rotdens", "pimc_add_pot3d",
                 "pimc_add_rot3d"):
        assert name in und, name
