"""CPU oracle for the HTFETI hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy/scipy FP64 restatement of the reference's method
(`ostravaTokyo/hfls`, files cited per function as ``path:line``), plus a C++
restatement of its PARDISO path for the CPU baseline (``cpu/``).  It exists so
that tests, the smoke check and the CPU-baseline leg of the benchmark can check
the HIP product path.  Nothing in the product (``hfls_amd``, ``include``) may
import, call or link it, and it is never the thing measured as the GPU result.

Modules
-------
mesh      structured cube mesh, numbering, Dirichlet/corner DOFs   (Mesh.cpp:69-289)
fe        hex8 solid45 element, CSR pattern, assembly               (Data.cpp:25-490)
decomp    l2g, interfaces, Bf / Bc / weights                        (Data.cpp:111-144,
                                                                     Cluster.cpp:682-959)
kernel    kernel detection, Gram-Schmidt, null pivots               (Matrix.cpp:835-884,
                                                                     1029-1105, 1976-2472)
feti      HTFETI operators, PCPG, recovery                          (Cluster.cpp:21-1458,
                                                                     1648-1678)
direct    global FP64 direct solve used as a brute-force pin
loads     seeded smoothed face tractions (the multi-load-case inputs)
cpu_port  C++ FP64 HTFETI (cpu/htfeti_cpu.cpp, OpenMP over subdomains): the
          reference's PARDISO path with analytic kernel and Dirichlet_implicit,
          band Cholesky instead of PARDISO; timed as the bench's CPU baseline,
          pinned against this package and the golden fixtures
          (tests/test_cpu_port.py)

Parity status per function is listed in each module's header.
"""
