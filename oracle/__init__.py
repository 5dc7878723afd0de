"""CPU oracle for the interferogram-formation hot path (test infrastructure).

This package is a restatement of the reference algorithms in FP64 NumPy (and a
C/OpenMP port in ``oracle/c``).  It is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import, call or link it.  The product (``sentinel_gpu_processor_amd``) never
routes through it.

Modules and the reference code each follows (file:line under the reference
repository):

- ``geom``      orbit.cu:15-105,170-279; bounds.cu:237-317; geo2rdr.cu:48-59,
                138-158,170-248,465-549
- ``ifg``       sario.cu:199-286; crossmul_daemon.cu:606-666;
                coherence.py:14-147; sario.py:397-420; interfere.py:16-180
- ``goldstein`` goldstein.cu:42-408
- ``sinc``      windowed-sinc interpolator (no reference counterpart)
- ``cport``     ctypes binding of ``oracle/c/s1cpu.c`` (built by
                ``make -C oracle all`` into ``oracle/libs1cpu.so``): the same
                chain in C11 + OpenMP, the CPU timing beside the GPU; its
                header cites the reference lines per function, and
                tests/test_oracle_cport.py pins it against the modules above

Parity status per function is listed in each module's header.
"""
