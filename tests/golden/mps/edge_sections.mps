* comment line, then a model that exercises every section
NAME          EDGES
OBJSENSE
    MAX
ROWS
 N  COST
 N  SPARE
 L  LIM1
 G  LIM2
 E  BAL
 E  BAL2
 L  LIM1
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    X1        COST         1.0   LIM1         1.0
    X1        LIM2         2.0   SPARE        9.0
    MARKER                 'MARKER'                 'INTEND'
    X2        COST        -2.0   LIM1         1.0
    X2        BAL          1.0   LIM1         0.5
    X3        COST         3.0   BAL2         1.0
    X3        LIM2         0.0
    X1        BAL          4.0
RHS
    RHS       COST        -7.5   LIM1         4.0
    RHS       LIM2         1.0   BAL          2.0
    RHS       BAL          9.0
    BAL2      3.0
RANGES
    RNG       LIM1         2.5   LIM2        -1.5
    RNG       BAL         -3.0   BAL2         2.0
BOUNDS
 UP BND       X1           4.0
 MI BND       X2
 UP BND       X2           8.0
 FR BND       X3
 PL BND       X3
 LI BND       X4           1.0
 UI BND       X5           2e30
 FX BND       X6           -1e31
 up BND       X3           5.0
ENDATA
