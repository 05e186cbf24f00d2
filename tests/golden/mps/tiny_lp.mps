NAME          TINYLP
ROWS
 N  COST
 L  R1
 L  R2
COLUMNS
    X         COST      -3.0       R1        1.0
    X         R2        1.0
    Y         COST      -2.0       R1        1.0
    Y         R2        3.0
RHS
    RHS       R1        4.0        R2        6.0
BOUNDS
 UP BND       X         3.0
ENDATA
