"""The ISO_C_BINDING module (fall3d_amd/fortran/f3d_mod.f90) against the C
header: every f3d_* entry point of include/fall3d_amd.h has a public bind(C)
interface of the same name and argument count, and a Fortran program that
calls each of them compiles with amdflang and links against
libfall3d_amd.so (no call is executed: this needs no GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "fall3d_amd.h")
MOD = os.path.join(ROOT, "fall3d_amd", "fortran", "f3d_mod.f90")
LIBDIR = os.path.join(ROOT, "fall3d_amd")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def header_entries():
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|void|long long|const char \*)\s*\*?\s*(f3d_\w+)\s*\(([^)]*)\)",
                         text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def module_entries():
    text = open(MOD).read()
    out = {}
    for m in re.finditer(r"(?:function|subroutine)\s+(\w+)\s*\(([^)]*)\)\s*&?\s*(?:\n\s*)?"
                         r"(?:[^\n]*?)bind\(C,\s*name='(f3d_\w+)'\)", text):
        args = m.group(2).replace("&", " ").strip()
        out[m.group(3)] = (m.group(1), 0 if not args else args.count(",") + 1)
    public = set()
    for m in re.finditer(r"public\s*::((?:[^\n]*&\s*\n)*[^\n]*)", text):
        public |= {w.strip() for w in m.group(1).replace("&", " ").split(",") if w.strip()}
    return out, public


def test_every_header_entry_has_a_public_binding():
    hdr = header_entries()
    mod, public = module_entries()
    assert len(hdr) >= 45
    missing = sorted(set(hdr) - set(mod))
    assert not missing, missing
    for name, nargs in hdr.items():
        fname, fargs = mod[name]
        assert fargs == nargs, (name, nargs, fargs)
        if name != "f3d_last_error":  # wrapped by f3d_error_message
            assert fname in public, fname


def _call_program():
    _, public = module_entries()
    mod, _ = module_entries()
    lines = ["program abi_check", "  use iso_c_binding", "  use fall3d_amd", "  implicit none",
             "  type(c_ptr) :: ctx, reg, batch, ctxs(1)", "  type(f3d_grid) :: grid",
             "  type(f3d_run_info) :: info", "  type(f3d_gravity), target :: gc",
             "  real(c_double) :: a(9), s, t, u, v", "  integer(c_int) :: rc, i, j, k, ia(2)",
             "  integer(c_long) :: n", "  integer(c_long_long) :: b1, b2",
             "  character(len=512) :: msg",
             "  if (command_argument_count() < 99) stop 0",
             "  rc = f3d_create(ctx, grid, F3D_FP32)",
             "  rc = f3d_create_ensemble(ctx, grid, 2)",
             "  rc = f3d_set_classes(ctx, a, a, a, 2)",
             "  rc = f3d_set_models(ctx, 0, 1, s, t, u)",
             "  rc = f3d_set_meteo(ctx, a, a, a, a, a, a, a, a, a, a)",
             "  rc = f3d_set_meteo_window(ctx, reg, 0, 0)",
             "  rc = f3d_get_dt(ctx, s, t)", "  rc = f3d_set_dt(ctx, s)",
             "  rc = f3d_set_source(ctx, 1, ia, ia, ia, a, s)",
             "  rc = f3d_set_member_source(ctx, 0, 1, ia, ia, ia, a, s)",
             "  rc = f3d_step(ctx, 1)", "  rc = f3d_budget(ctx, a)",
             "  rc = f3d_ensemble_budget(ctx, a)", "  rc = f3d_get_state(ctx, a, a)",
             "  rc = f3d_get_member_state(ctx, 0, a, a)", "  rc = f3d_set_state(ctx, a, a)",
             "  rc = f3d_get_coef(ctx, a, a, a, a, a, a)", "  rc = f3d_set_landuse(ctx, a, a)",
             "  rc = f3d_set_topography(ctx, a)", "  rc = f3d_set_precipitation(ctx, a)",
             "  rc = f3d_set_gravity_current(ctx, c_loc(gc), 1, a)",
             "  rc = f3d_get_gravity_current(ctx, a, s, i)", "  rc = f3d_get_divu(ctx, a)",
             "  rc = f3d_output_products(ctx, 0, a, a, a)", "  rc = f3d_get_vdep(ctx, a)",
             "  rc = f3d_set_pressure(ctx, a)",
             "  rc = f3d_output_concentration(ctx, 0, a, a, 0, 1, a, a)",
             "  rc = f3d_write_restart(ctx, 0, 'r.nc' // c_null_char, info)",
             "  rc = f3d_read_restart(ctx, 'r.nc' // c_null_char, a, info)",
             "  rc = f3d_restart_write('r.nc' // c_null_char, 2, 2, 2, 1, 0, a, a, info)",
             "  rc = f3d_restart_read('r.nc' // c_null_char, 2, 2, 2, 1, a, a, info)",
             "  rc = f3d_get_clock(ctx, s, i, t, u)", "  rc = f3d_set_clock(ctx, s, 1, t, u)",
             "  rc = f3d_get_cd_state(ctx, s)", "  rc = f3d_set_cd_state(ctx, s)",
             "  rc = f3d_sync(ctx)", "  rc = f3d_set_tile(ctx, 58, 18)",
             "  rc = f3d_set_tile_config(ctx, 58, 18, 0, 2)", "  rc = f3d_timing(ctx, 1)",
             "  rc = f3d_kernel_time(ctx, 1, s, n)", "  b1 = f3d_device_bytes(ctx)",
             "  rc = f3d_autotune(ctx, 2, i, j, k, s)",
             "  rc = f3d_batch_create(batch, ctxs, 1)", "  rc = f3d_batch_set_tile(batch, 58, 18, 2)",
             "  rc = f3d_batch_get_tile(batch, i, j, k)", "  rc = f3d_batch_use_graph(batch, 1)",
             "  rc = f3d_batch_step(batch, 1)", "  rc = f3d_batch_sync(batch)",
             "  rc = f3d_batch_timing(batch, 1)", "  rc = f3d_batch_kernel_time(batch, 1, s, n)",
             "  rc = f3d_batch_autotune(batch, 2, i, j, k, s)", "  call f3d_batch_destroy(batch)",
             "  rc = f3d_mem_info(b1, b2)", "  rc = f3d_hbm_copy(b1, 1, s)",
             "  rc = f3d_cmass3d_noalloc(ctx, ctx, ctx, ctx, 2, 2, 2, 1, 1, s, t, ctx, ctx, ctx)",
             "  msg = f3d_error_message()", "  call f3d_destroy(ctx)",
             "  print *, rc, msg, v", "end program abi_check"]
    called = {m.group(1) for m in re.finditer(r"(f3d_\w+)\(", "\n".join(lines))}
    return "\n".join(lines) + "\n", called


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not in this image")
def test_fortran_program_calling_every_binding_links(tmp_path):
    lib = os.path.join(LIBDIR, "libfall3d_amd.so")
    if not os.path.exists(lib):
        pytest.skip("libfall3d_amd.so not built")
    src, called = _call_program()
    mod, _ = module_entries()
    assert {f for f, _ in mod.values()} - {"f3d_last_error_c"} <= called | {"f3d_error_message"}
    prog = tmp_path / "abi_check.f90"
    prog.write_text(src)
    out = tmp_path / "abi_check"
    r = subprocess.run([FLANG, "-O0", "-module-dir", str(tmp_path), MOD, str(prog),
                        "-L" + LIBDIR, "-lfall3d_amd", "-Wl,-rpath," + LIBDIR, "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert out.exists()
