"""Problem generators: initial conditions for the configurations the library is
measured and checked on.  Each returns per-block AthenaArray-layout host arrays
(``u``, ``b1``, ``b2``, ``b3`` and, for SR problems, ``w``/``w1``)."""
