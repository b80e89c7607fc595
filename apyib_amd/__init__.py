"""apyib_amd: MI355X-native correlated-wavefunction hot path of apyib.

The wavefunction classes keep apyib's call signatures (ci_wfn, mp2_wfn,
energy, utils); their contractions run in the HIP library libapyib_amd.so
through the C ABI declared in include/apyib_amd.h.
"""
__version__ = "0.1.0"
