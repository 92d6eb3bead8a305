"""MI355X (gfx950) implementation of one CReSS large time step.

cress_amd.host   -- the C ABI (include/cress_hip.h) from Python: Solver
cress_amd.cases  -- grids, base states and initial states of the workloads
cress_amd.build  -- builds cress_amd/_lib/libcresship.so with hipcc
"""
