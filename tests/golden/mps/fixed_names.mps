NAME          FIXEDNM
ROWS
 N  COST
 G  ROW A
 L  ROW B
COLUMNS
    COL 1     COST         1.0         ROW A        1.0
    COL 1     ROW B        2.0
    COL 2     COST         2.0         ROW A        1.0
RHS
    RHS       ROW A        1.0         ROW B        6.0
BOUNDS
 UP BND       COL 2        3.0
ENDATA
