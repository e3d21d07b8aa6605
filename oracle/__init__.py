"""CPU oracle of the spectral random-field path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; the product (randomfield_amd, librf_hip.so) never does.
"""
