NAME          WOODINFE
ROWS
 N  COST
 E  SPIDE
 E  DPIDE
 E  SPICH
 E  DPICH
 E  SPISF
 E  DPISF
 E  SPILA
 E  DPILA
 E  SPISE
 E  DPISE
 E  DWADE
 E  SWACH
 E  DWACH
 E  SWASF
 E  DWASF
 E  DWALA
 E  SWASE
 E  DWASE
 E  SOKDE
 E  DOKDE
 E  SOKCH
 E  DOKCH
 E  SOKSF
 E  DOKSF
 E  SOKLA
 E  DOKLA
 E  SOKSE
 E  DOKSE
 E  DMODE
 E  DMOCH
 E  SMOSF
 E  DMOSF
 E  DMOLA
 E  SMOSE
 E  DMOSE
COLUMNS
    SPIDE1    SPIDE               1.   COST               10.
    SPICH1    SPICH               1.   COST               11.
    SPISF1    SPISF               1.   COST               12.
    SPILA1    SPILA               1.   COST               15.
    SPISE1    SPISE               1.   COST               11.
    SPIDE2    SPIDE               1.   COST               12.
    SPICH2    SPICH               1.   COST               15.
    SPISF2    SPISF               1.   COST               12.
    SWACH1    SWACH               1.   COST               50.
    SWASF1    SWASF               1.   COST               50.
    SWASE1    SWASE               1.   COST               60.
    SOKDE1    SOKDE               1.   COST               30.
    SOKCH1    SOKCH               1.   COST               35.
    SOKSF1    SOKSF               1.   COST               40.
    SOKLA1    SOKLA               1.   COST               40.
    SOKSE1    SOKSE               1.   COST               50.
    SMOSF1    SMOSF               1.   COST               45.
    SMOSE1    SMOSE               1.   COST               55.
    DPIDE1    DPIDE              -1.
    DPICH1    DPICH              -1.
    DPISF1    DPISF              -1.
    DPILA1    DPILA              -1.
    DPISE1    DPISE              -1.
    DWADE1    DWADE              -1.
    DWACH1    DWACH              -1.
    DWASF1    DWASF              -1.
    DWALA1    DWALA              -1.
    DWASE1    DWASE              -1.
    DOKDE1    DOKDE              -1.
    DOKCH1    DOKCH              -1.
    DOKSF1    DOKSF              -1.
    DOKLA1    DOKLA              -1.
    DOKSE1    DOKSE              -1.
    DMODE1    DMODE              -1.
    DMOCH1    DMOCH              -1.
    DMOSF1    DMOSF              -1.
    DMOLA1    DMOLA              -1.
    DMOSE1    DMOSE              -1.
    TPIDECH   COST               10.   SPIDE              -1.
    TPIDECH   DPICH               1.
    TOKDECH   COST               10.   SOKDE              -1.
    TOKDECH   DOKCH               1.
    TPIDESF   COST               12.   SPIDE              -1.
    TPIDESF   DPISF               1.
    TOKDESF   COST               12.   SOKDE              -1.
    TOKDESF   DOKSF               1.
    TPIDELA   COST               12.   SPIDE              -1.
    TPIDELA   DPILA               1.
    TOKDELA   COST               12.   SOKDE              -1.
    TOKDELA   DOKLA               1.
    TPIDESE   COST               15.   SPIDE              -1.
    TPIDESE   DPISE               1.
    TOKDESE   COST               15.   SOKDE              -1.
    TOKDESE   DOKSE               1.
    TPICHDE   COST               11.   SPICH              -1.
    TPICHDE   DPIDE               1.
    TWACHDE   COST               11.   SWACH              -1.
    TWACHDE   DWADE               1.
    TOKCHDE   COST               11.   SOKCH              -1.
    TOKCHDE   DOKDE               1.
    TPICHSF   COST               15.   SPICH              -1.
    TPICHSF   DPISF               1.
    TWACHSF   COST               15.   SWACH              -1.
    TWACHSF   DWASF               1.
    TOKCHSF   COST               15.   SOKCH              -1.
    TOKCHSF   DOKSF               1.
    TPICHLA   COST               18.   SPICH              -1.
    TPICHLA   DPILA               1.
    TWACHLA   COST               18.   SWACH              -1.
    TWACHLA   DWALA               1.
    TOKCHLA   COST               18.   SOKCH              -1.
    TOKCHLA   DOKLA               1.
    TPISFDE   COST               12.   SPISF              -1.
    TPISFDE   DPIDE               1.
    TWASFDE   COST               12.   SWASF              -1.
    TWASFDE   DWADE               1.
    TOKSFDE   COST               12.   SOKSF              -1.
    TOKSFDE   DOKDE               1.
    TMOSFDE   COST               12.   SMOSF              -1.
    TMOSFDE   DMODE               1.
    TPISFCH   COST               15.   SPISF              -1.
    TPISFCH   DPICH               1.
    TWASFCH   COST               15.   SWASF              -1.
    TWASFCH   DWACH               1.
    TOKSFCH   COST               15.   SOKSF              -1.
    TOKSFCH   DOKCH               1.
    TMOSFCH   COST               15.   SMOSF              -1.
    TMOSFCH   DMOCH               1.
    TPISFLA   COST                5.   SPISF              -1.
    TPISFLA   DPILA               1.
    TWASFLA   COST                5.   SWASF              -1.
    TWASFLA   DWALA               1.
    TOKSFLA   COST                5.   SOKSF              -1.
    TOKSFLA   DOKLA               1.
    TMOSFLA   COST                5.   SMOSF              -1.
    TMOSFLA   DMOLA               1.
    TPILADE   COST               14.   SPILA              -1.
    TPILADE   DPIDE               1.
    TOKLADE   COST               14.   SOKLA              -1.
    TOKLADE   DOKDE               1.
    TPILACH   COST               16.   SPILA              -1.
    TPILACH   DPICH               1.
    TOKLACH   COST               16.   SOKLA              -1.
    TOKLACH   DOKCH               1.
    TPILASF   COST                5.   SPILA              -1.
    TPILASF   DPISF               1.
    TOKLASF   COST                5.   SOKLA              -1.
    TOKLASF   DOKSF               1.
    TPISEDE   COST               20.   SPISE              -1.
    TPISEDE   DPIDE               1.
    TWASEDE   COST               20.   SWASE              -1.
    TWASEDE   DWADE               1.
    TOKSEDE   COST               20.   SOKSE              -1.
    TOKSEDE   DOKDE               1.
    TMOSEDE   COST               20.   SMOSE              -1.
    TMOSEDE   DMODE               1.
    TPISECH   COST               18.   SPISE              -1.
    TPISECH   DPICH               1.
    TWASECH   COST               18.   SWASE              -1.
    TWASECH   DWACH               1.
    TOKSECH   COST               18.   SOKSE              -1.
    TOKSECH   DOKCH               1.
    TMOSECH   COST               18.   SMOSE              -1.
    TMOSECH   DMOCH               1.
    TPISESF   COST               10.   SPISE              -1.
    TPISESF   DPISF               1.
    TWASESF   COST               10.   SWASE              -1.
    TWASESF   DWASF               1.
    TOKSESF   COST               10.   SOKSE              -1.
    TOKSESF   DOKSF               1.
    TMOSESF   COST               10.   SMOSE              -1.
    TMOSESF   DMOSF               1.
    TPISELA   COST               15.   SPISE              -1.
    TPISELA   DPILA               1.
    TWASELA   COST               15.   SWASE              -1.
    TWASELA   DWALA               1.
    TOKSELA   COST               15.   SOKSE              -1.
    TOKSELA   DOKLA               1.
    TMOSELA   COST               15.   SMOSE              -1.
    TMOSELA   DMOLA               1.
RHS
BOUNDS
 UP BNDSET    SPIDE1            100.
 UP BNDSET    SPICH1            100.
 UP BNDSET    SPISF1             90.
 UP BNDSET    SPILA1             50.
 UP BNDSET    SPISE1             10.
 UP BNDSET    SPIDE2            200.
 UP BNDSET    SPICH2            100.
 UP BNDSET    SWACH1             25.
 UP BNDSET    SWASF1             10.
 UP BNDSET    SOKDE1             50.
 UP BNDSET    SOKCH1             40.
 UP BNDSET    SOKSF1             20.
 UP BNDSET    SOKLA1              5.
 UP BNDSET    SMOSF1             30.
 LO BNDSET    DPIDE1            100.
 LO BNDSET    DPICH1            100.
 LO BNDSET    DPISF1             90.
 LO BNDSET    DPILA1             50.
 LO BNDSET    DPISE1             10.
 LO BNDSET    DWADE1             20.
 LO BNDSET    DWACH1             25.
 LO BNDSET    DWASF1             10.
 LO BNDSET    DWALA1             15.
 LO BNDSET    DWASE1              5.
 LO BNDSET    DOKDE1             50.
 LO BNDSET    DOKCH1             40.
 LO BNDSET    DOKSF1             20.
 LO BNDSET    DOKLA1              5.
 LO BNDSET    DOKSE1             15.
 LO BNDSET    DMODE1             20.
 LO BNDSET    DMOCH1             25.
 LO BNDSET    DMOSF1             30.
 LO BNDSET    DMOLA1             20.
 LO BNDSET    DMOSE1             10.
ENDATA
