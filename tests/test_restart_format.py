"""ccq's amplitude files are Fortran unformatted sequential records
(src/diis.F90:329-345, src/utils/cc_utils.f90:89-145).  The library's record
reader / writer (ccq_amd/csrc/restart.cpp) must interoperate with a real
Fortran runtime: a file written by a Fortran WRITE is read back exactly, and
a file the library writes is read by a Fortran READ.  Host code only."""

import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/bin/amdflang"

FORTRAN = """
program rec
    implicit none
    integer :: n, i
    real(kind=8), allocatable :: x(:)
    character(len=16) :: mode
    call get_command_argument(1, mode)
    n = 1000
    allocate(x(n))
    if (trim(mode) == 'write') then
        do i = 1, n
            x(i) = 1.0d0 / i - 0.25d0 * i
        end do
        open(10, file='f.bin', form='unformatted', status='unknown')
        write(10) x
        close(10)
    else
        open(10, file='c.bin', form='unformatted', status='old')
        read(10) x
        close(10)
        do i = 1, n
            if (x(i) /= 3.0d0 * i - 1.0d0 / (i + 1)) stop 3
        end do
        print '(a)', 'read-ok'
    end if
end program rec
"""

DRIVER = r"""
#include <cstdio>
#include <string>
#include <vector>
#include "restart_io.h"
int main(int argc, char** argv) {
  const int n = 1000;
  std::vector<double> x(n);
  if (std::string(argv[1]) == "write") {
    for (int i = 1; i <= n; ++i) x[i - 1] = 3.0 * i - 1.0 / (i + 1);
    return ccqio::write_record("c.bin", x.data(), n) ? 0 : 1;
  }
  std::string why;
  if (!ccqio::read_record("f.bin", x.data(), n, why)) { std::puts(why.c_str()); return 1; }
  for (int i = 1; i <= n; ++i) if (x[i - 1] != 1.0 / i - 0.25 * i) return 2;
  std::vector<double> y(n + 1);
  if (ccqio::read_record("f.bin", y.data(), n + 1, why)) return 4;   // wrong length refused
  std::puts("read-ok");
  return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(FC) or not shutil.which("g++"),
                    reason="needs amdflang and g++")
def test_records_interoperate_with_fortran(tmp_path):
    (tmp_path / "rec.f90").write_text(FORTRAN)
    (tmp_path / "drv.cpp").write_text(DRIVER)
    src = os.path.join(ROOT, "ccq_amd", "csrc")
    subprocess.run([FC, "-O1", "-fdefault-integer-8", "rec.f90", "-o", "rec"], cwd=tmp_path,
                   check=True)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", src, "drv.cpp",
                    os.path.join(src, "restart.cpp"), "-o", "drv"], cwd=tmp_path, check=True)
    subprocess.run(["./rec", "write"], cwd=tmp_path, check=True)
    out = subprocess.run(["./drv", "read"], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode == 0 and "read-ok" in out.stdout, out.stdout
    subprocess.run(["./drv", "write"], cwd=tmp_path, check=True)
    out = subprocess.run(["./rec", "read"], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode == 0 and "read-ok" in out.stdout, out.stdout + out.stderr
    # the record layout itself: 4-byte length, data, 4-byte length
    raw = (tmp_path / "c.bin").read_bytes()
    assert len(raw) == 8000 + 8
    assert np.frombuffer(raw[:4], np.int32)[0] == 8000 == np.frombuffer(raw[-4:], np.int32)[0]
