"""Oracle: CPU restatements of the reference's algorithms for the M-plicits hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this package, and only as the checker (or
as the timed CPU baseline).  The product package `mplicits` never imports it
and has no CPU path of its own.

Modules (each function cites the reference file:line it restates):
  siren    - SIREN forward / input gradient / init call order (FP64 numpy, torch)
  grid     - gen_mc_coordinate_grid as torch evaluates it (float32 numpy)
  extract  - create_mesh_multistage's chunked band-culled sampling loop
  losses   - loss_add_detail_fine and the benchmark's L1 / eikonal / cosine loss
             (torch FP64 autograd, double backward)
  mc       - Lewiner marching cubes (C, oracle/mc_lewiner.c, built by build.py)
  trace    - multiscale sphere-tracing step rule and ray generation
  render   - the renderer's fixed-iteration frame and Phong shading
  meshdist - point-to-mesh distance, calc_deltas, neighbourhood sampling
"""
