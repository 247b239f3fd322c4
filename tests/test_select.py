"""The device/host restatement of std::nth_element (treecorr_amd/csrc/tc_select.h)
leaves exactly the permutation libstdc++ leaves (checked on the host)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_nth_element_matches_libstdcxx(tmp_path):
    exe = str(tmp_path / 'select_check')
    subprocess.check_call(['g++', '-O2', '-std=c++17', os.path.join(HERE, 'native', 'select_check.cpp'),
                           '-o', exe])
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert '0 of' in out.stdout
