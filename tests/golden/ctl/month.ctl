DSET    ^month.dat                                                                                                                       
TITLE    monthly means
UNDEF    0.10000E+21  ! gap value
XDEF     100  LINEAR    -180.00000          -180.00000    
YDEF      50  LINEAR    -90.000000          -90.000000    
ZDEF       1  LINEAR     0.0000000           1.0000000    
TDEF       3  LINEAR  12:00Z29FEB1980        1mo
OPTIONS  365_DAY_CALENDAR
VARS 1  ! Number of variables
tmean            1  1 VARIABLES 
ENDVARS

