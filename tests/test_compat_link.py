"""The drop-in link of the reference's own host files against
libdgrain_compat.so (CPU only: link, no run)."""
import os

import pytest


@pytest.mark.ref
def test_reference_host_files_link_unchanged():
    """The drop-in link: the reference's own main.c, extract_mesh_attributes.c,
    initialize.c and boundary_conditions.c (2DCode), compiled unchanged with
    the reference's flags and linked against libdgrain_compat.so in place of
    its operator files (oracle/Makefile, target compat-link; no unresolved
    symbol allowed).  Every symbol the reference objects leave undefined,
    apart from libc and libm, is one the shim exports.  Link only: running
    the program needs a GPU."""
    import subprocess
    from conftest import require_ref
    require_ref()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["make", "-s", "-C", os.path.join(root, "oracle"), "compat-link"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    exe = os.path.join(root, "oracle", "_ref", "dgrain_2dcode")
    assert os.access(exe, os.X_OK)
    undef = subprocess.run(["nm", "-u", exe], capture_output=True, text=True, check=True).stdout.split()
    undef = {u.split("@")[0] for u in undef if u not in ("U", "w")}
    libc = {u for u in undef if u.startswith("_") or u in ("printf", "fopen", "fclose", "fscanf", "fgets", "malloc",
                                                           "calloc", "free", "exit", "strcpy", "strcat", "sprintf",
                                                           "puts", "putchar", "fprintf", "fflush", "sqrt", "pow",
                                                           "fabs", "memset", "memcpy", "strlen", "atoi", "atof",
                                                           "strcmp", "sscanf", "fwrite", "fread", "perror")}
    shim = subprocess.run(["nm", "-D", "--defined-only", os.path.join(root, "dg-rain_amd", "lib", "libdgrain_compat.so")],
                          capture_output=True, text=True, check=True).stdout.split()
    missing = sorted(undef - libc - set(shim))
    assert not missing, missing
    assert "time_evolution" in undef
