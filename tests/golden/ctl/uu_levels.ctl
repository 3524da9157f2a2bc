DSET    ^uu_levels.dat                                                                                                                   
TITLE    zonal velocity, m/s
UNDEF   -0.10000E+33  ! gap value
XDEF     286  LINEAR     27.500000          0.50000000E-01
YDEF     159  LINEAR     40.940000          0.40000000E-01
ZDEF      20  LEVELS    2.5000000       54.868421       107.23684       159.60526       211.97368       264.34211       316.71053    
                        369.07895       421.44737       473.81579       526.18421       578.55263       630.92105       683.28947    
                        735.65789       788.02632       840.39474       892.76316       945.13158       997.50000    
TDEF      12  LINEAR  23:59Z31DEC2001        1hr
OPTIONS  365_DAY_CALENDAR
VARS 1  ! Number of variables
u           20  1 VARIABLES 
ENDVARS

