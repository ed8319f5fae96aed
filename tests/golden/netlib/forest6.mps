NAME          FOREST
ROWS
 N  COST
 G  D1
 E  BBR1
 G  WBR1
 E  BSM1
 G  WSM1
 E  BBB1
 G  WBB1
 E  BNR1
 G  WNR1
 E  BGA1
 G  WGA1
 G  D2
 E  BBR2
 G  WBR2
 E  BSM2
 G  WSM2
 E  BBB2
 G  WBB2
 E  BNR2
 G  WNR2
 E  BGA2
 G  WGA2
 G  D3
 E  BBR3
 G  WBR3
 E  BSM3
 G  WSM3
 E  BBB3
 G  WBB3
 E  BNR3
 G  WNR3
 E  BGA3
 G  WGA3
 G  D4
 E  BBR4
 G  WBR4
 E  BSM4
 G  WSM4
 E  BBB4
 G  WBB4
 E  BNR4
 G  WNR4
 E  BGA4
 G  WGA4
 G  D5
 E  BBR5
 G  WBR5
 E  BSM5
 G  WSM5
 E  BBB5
 G  WBB5
 E  BNR5
 G  WNR5
 E  BGA5
 G  WGA5
 G  D6
 E  BBR6
 G  WBR6
 E  BSM6
 G  WSM6
 E  BBB6
 G  WBB6
 E  BNR6
 G  WNR6
 E  BGA6
 G  WGA6
COLUMNS
    IBR0      BBR1                1.   WBR1              -.75
    ISM0      BSM1                1.   WSM1               -.5
    IBB0      BBB1                1.   WBB1              -.25
    INR0      BNR1                1.   WNR1               -.3
    IGA0      BGA1                1.   WGA1               -.5
    IBR1      BBR1               -1.   BBR2                1.
    IBR1      WBR2              -.75
    HBR1      BBR1               -1.   D1                  .4
    HBR1      COST                9.
    WBR1      BBR1               -1.   WBR1                1.
    WBR1      COST               4.5
    ISM1      BSM1               -1.   BSM2                1.
    ISM1      WSM2               -.5
    HSM1      BSM1               -1.   D1                  .2
    HSM1      COST              13.5
    WSM1      BSM1               -1.   WSM1                1.
    WSM1      COST               1.8
    IBB1      BBB1               -1.   BBB2                1.
    IBB1      WBB2              -.25
    HBB1      BBB1               -1.   D1                  .3
    HBB1      COST                9.
    WBB1      BBB1               -1.   WBB1                1.
    WBB1      COST               2.7
    INR1      BNR1               -1.   BNR2                1.
    INR1      WNR2               -.3
    HNR1      BNR1               -1.   D1                  .4
    HNR1      COST                9.
    WNR1      BNR1               -1.   WNR1                1.
    WNR1      COST               3.6
    IGA1      BGA1               -1.   BGA2                1.
    IGA1      WGA2               -.5
    HGA1      BGA1               -1.   D1                  .3
    HGA1      COST              13.5
    WGA1      BGA1               -1.   WGA1                1.
    WGA1      COST               1.8
    IBR2      BBR2               -1.   BBR3                1.
    IBR2      WBR3              -.75
    HBR2      BBR2               -1.   D2                  .4
    HBR2      COST               8.1
    WBR2      BBR2               -1.   WBR2                1.
    WBR2      COST              4.05
    ISM2      BSM2               -1.   BSM3                1.
    ISM2      WSM3               -.5
    HSM2      BSM2               -1.   D2                  .2
    HSM2      COST             12.15
    WSM2      BSM2               -1.   WSM2                1.
    WSM2      COST              1.62
    IBB2      BBB2               -1.   BBB3                1.
    IBB2      WBB3              -.25
    HBB2      BBB2               -1.   D2                  .3
    HBB2      COST               8.1
    WBB2      BBB2               -1.   WBB2                1.
    WBB2      COST              2.43
    INR2      BNR2               -1.   BNR3                1.
    INR2      WNR3               -.3
    HNR2      BNR2               -1.   D2                  .4
    HNR2      COST               8.1
    WNR2      BNR2               -1.   WNR2                1.
    WNR2      COST              3.24
    IGA2      BGA2               -1.   BGA3                1.
    IGA2      WGA3               -.5
    HGA2      BGA2               -1.   D2                  .3
    HGA2      COST             12.15
    WGA2      BGA2               -1.   WGA2                1.
    WGA2      COST              1.62
    IBR3      BBR3               -1.   BBR4                1.
    IBR3      WBR4              -.75
    HBR3      BBR3               -1.   D3                  .4
    HBR3      COST          7.289999
    WBR3      BBR3               -1.   WBR3                1.
    WBR3      COST             3.645
    ISM3      BSM3               -1.   BSM4                1.
    ISM3      WSM4               -.5
    HSM3      BSM3               -1.   D3                  .2
    HSM3      COST            10.935
    WSM3      BSM3               -1.   WSM3                1.
    WSM3      COST             1.458
    IBB3      BBB3               -1.   BBB4                1.
    IBB3      WBB4              -.25
    HBB3      BBB3               -1.   D3                  .3
    HBB3      COST          7.289999
    WBB3      BBB3               -1.   WBB3                1.
    WBB3      COST             2.187
    INR3      BNR3               -1.   BNR4                1.
    INR3      WNR4               -.3
    HNR3      BNR3               -1.   D3                  .4
    HNR3      COST          7.289999
    WNR3      BNR3               -1.   WNR3                1.
    WNR3      COST             2.916
    IGA3      BGA3               -1.   BGA4                1.
    IGA3      WGA4               -.5
    HGA3      BGA3               -1.   D3                  .3
    HGA3      COST            10.935
    WGA3      BGA3               -1.   WGA3                1.
    WGA3      COST             1.458
    IBR4      BBR4               -1.   BBR5                1.
    IBR4      WBR5              -.75
    HBR4      BBR4               -1.   D4                  .4
    HBR4      COST          6.560999
    WBR4      BBR4               -1.   WBR4                1.
    WBR4      COST            3.2805
    ISM4      BSM4               -1.   BSM5                1.
    ISM4      WSM5               -.5
    HSM4      BSM4               -1.   D4                  .2
    HSM4      COST          9.841498
    WSM4      BSM4               -1.   WSM4                1.
    WSM4      COST            1.3122
    IBB4      BBB4               -1.   BBB5                1.
    IBB4      WBB5              -.25
    HBB4      BBB4               -1.   D4                  .3
    HBB4      COST          6.560999
    WBB4      BBB4               -1.   WBB4                1.
    WBB4      COST            1.9683
    INR4      BNR4               -1.   BNR5                1.
    INR4      WNR5               -.3
    HNR4      BNR4               -1.   D4                  .4
    HNR4      COST          6.560999
    WNR4      BNR4               -1.   WNR4                1.
    WNR4      COST            2.6244
    IGA4      BGA4               -1.   BGA5                1.
    IGA4      WGA5               -.5
    HGA4      BGA4               -1.   D4                  .3
    HGA4      COST          9.841498
    WGA4      BGA4               -1.   WGA4                1.
    WGA4      COST            1.3122
    IBR5      BBR5               -1.   BBR6                1.
    IBR5      WBR6              -.75
    HBR5      BBR5               -1.   D5                  .4
    HBR5      COST          5.904899
    WBR5      BBR5               -1.   WBR5                1.
    WBR5      COST           2.95245
    ISM5      BSM5               -1.   BSM6                1.
    ISM5      WSM6               -.5
    HSM5      BSM5               -1.   D5                  .2
    HSM5      COST          8.857349
    WSM5      BSM5               -1.   WSM5                1.
    WSM5      COST           1.18098
    IBB5      BBB5               -1.   BBB6                1.
    IBB5      WBB6              -.25
    HBB5      BBB5               -1.   D5                  .3
    HBB5      COST          5.904899
    WBB5      BBB5               -1.   WBB5                1.
    WBB5      COST           1.77147
    INR5      BNR5               -1.   BNR6                1.
    INR5      WNR6               -.3
    HNR5      BNR5               -1.   D5                  .4
    HNR5      COST          5.904899
    WNR5      BNR5               -1.   WNR5                1.
    WNR5      COST           2.36196
    IGA5      BGA5               -1.   BGA6                1.
    IGA5      WGA6               -.5
    HGA5      BGA5               -1.   D5                  .3
    HGA5      COST          8.857349
    WGA5      BGA5               -1.   WGA5                1.
    WGA5      COST           1.18098
    IBR6      BBR6               -1.
    HBR6      BBR6               -1.   D6                  .4
    HBR6      COST           5.31441
    WBR6      BBR6               -1.   WBR6                1.
    WBR6      COST          2.657205
    ISM6      BSM6               -1.
    HSM6      BSM6               -1.   D6                  .2
    HSM6      COST          7.971614
    WSM6      BSM6               -1.   WSM6                1.
    WSM6      COST          1.062882
    IBB6      BBB6               -1.
    HBB6      BBB6               -1.   D6                  .3
    HBB6      COST           5.31441
    WBB6      BBB6               -1.   WBB6                1.
    WBB6      COST          1.594323
    INR6      BNR6               -1.
    HNR6      BNR6               -1.   D6                  .4
    HNR6      COST           5.31441
    WNR6      BNR6               -1.   WNR6                1.
    WNR6      COST          2.125764
    IGA6      BGA6               -1.
    HGA6      BGA6               -1.   D6                  .3
    HGA6      COST          7.971614
    WGA6      BGA6               -1.   WGA6                1.
    WGA6      COST          1.062882
RHS
    RHS       D1               1050.   D2               1100.
    RHS       D3               1150.   D4               1200.
    RHS       D5               1250.   D6               1300.
BOUNDS
 UP LAND      IBR0            20000.
 UP LAND      ISM0            30000.
 UP LAND      IBB0            10000.
 UP LAND      INR0            25000.
 UP LAND      IGA0            50000.
ENDATA
