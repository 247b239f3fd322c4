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


def test_wrap_period_matches_reference_loop_and_is_bounded(tmp_path):
    """wrap_period (tc_common.h) is the reference's Position::wrap loop inside 1.5 periods
    and ends for tiny periods and far-out separations (no spinning kernel)."""
    exe = str(tmp_path / 'wrap_check')
    subprocess.check_call(['g++', '-O2', '-std=c++17', '-I', os.path.join(HERE, 'native', 'hip_emu'),
                           '-I', os.path.join(os.path.dirname(HERE), 'treecorr_amd', 'csrc'),
                           os.path.join(HERE, 'native', 'wrap_check.cpp'), '-o', exe, '-pthread'])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert 'in-box mismatches: 0 of' in out.stdout
