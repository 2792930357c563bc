NAME MRPL_CRUDE_8k_14d
OBJSENSE
    MAX
ROWS
 N obj
 E lift_0
 E lift_1
 E lift_2
 E lift_3
 E lift_4
 E lift_5
 E lift_6
 E lift_7
 L berth_0
 E tank_balance_arab_light_0
 E tank_balance_arab_heavy_0
 E tank_balance_basrah_medium_0
 E tank_balance_murban_0
 E tank_balance_upper_zakum_0
 E tank_balance_kuwait_export_0
 L cdu_max_0
 G cdu_min_0
 L cdu_sulfur_0
 L tank_farm_0
 L berth_1
 E tank_balance_arab_light_1
 E tank_balance_arab_heavy_1
 E tank_balance_basrah_medium_1
 E tank_balance_murban_1
 E tank_balance_upper_zakum_1
 E tank_balance_kuwait_export_1
 L cdu_max_1
 G cdu_min_1
 L cdu_sulfur_1
 L tank_farm_1
 L berth_2
 E tank_balance_arab_light_2
 E tank_balance_arab_heavy_2
 E tank_balance_basrah_medium_2
 E tank_balance_murban_2
 E tank_balance_upper_zakum_2
 E tank_balance_kuwait_export_2
 L cdu_max_2
 G cdu_min_2
 L cdu_sulfur_2
 L tank_farm_2
 L berth_3
 E tank_balance_arab_light_3
 E tank_balance_arab_heavy_3
 E tank_balance_basrah_medium_3
 E tank_balance_murban_3
 E tank_balance_upper_zakum_3
 E tank_balance_kuwait_export_3
 L cdu_max_3
 G cdu_min_3
 L cdu_sulfur_3
 L tank_farm_3
 L berth_4
 E tank_balance_arab_light_4
 E tank_balance_arab_heavy_4
 E tank_balance_basrah_medium_4
 E tank_balance_murban_4
 E tank_balance_upper_zakum_4
 E tank_balance_kuwait_export_4
 L cdu_max_4
 G cdu_min_4
 L cdu_sulfur_4
 L tank_farm_4
 L berth_5
 E tank_balance_arab_light_5
 E tank_balance_arab_heavy_5
 E tank_balance_basrah_medium_5
 E tank_balance_murban_5
 E tank_balance_upper_zakum_5
 E tank_balance_kuwait_export_5
 L cdu_max_5
 G cdu_min_5
 L cdu_sulfur_5
 L tank_farm_5
 L berth_6
 E tank_balance_arab_light_6
 E tank_balance_arab_heavy_6
 E tank_balance_basrah_medium_6
 E tank_balance_murban_6
 E tank_balance_upper_zakum_6
 E tank_balance_kuwait_export_6
 L cdu_max_6
 G cdu_min_6
 L cdu_sulfur_6
 L tank_farm_6
 L berth_7
 E tank_balance_arab_light_7
 E tank_balance_arab_heavy_7
 E tank_balance_basrah_medium_7
 E tank_balance_murban_7
 E tank_balance_upper_zakum_7
 E tank_balance_kuwait_export_7
 L cdu_max_7
 G cdu_min_7
 L cdu_sulfur_7
 L tank_farm_7
 L berth_8
 E tank_balance_arab_light_8
 E tank_balance_arab_heavy_8
 E tank_balance_basrah_medium_8
 E tank_balance_murban_8
 E tank_balance_upper_zakum_8
 E tank_balance_kuwait_export_8
 L cdu_max_8
 G cdu_min_8
 L cdu_sulfur_8
 L tank_farm_8
 L berth_9
 E tank_balance_arab_light_9
 E tank_balance_arab_heavy_9
 E tank_balance_basrah_medium_9
 E tank_balance_murban_9
 E tank_balance_upper_zakum_9
 E tank_balance_kuwait_export_9
 L cdu_max_9
 G cdu_min_9
 L cdu_sulfur_9
 L tank_farm_9
 L berth_10
 E tank_balance_arab_light_10
 E tank_balance_arab_heavy_10
 E tank_balance_basrah_medium_10
 E tank_balance_murban_10
 E tank_balance_upper_zakum_10
 E tank_balance_kuwait_export_10
 L cdu_max_10
 G cdu_min_10
 L cdu_sulfur_10
 L tank_farm_10
 L berth_11
 E tank_balance_arab_light_11
 E tank_balance_arab_heavy_11
 E tank_balance_basrah_medium_11
 E tank_balance_murban_11
 E tank_balance_upper_zakum_11
 E tank_balance_kuwait_export_11
 L cdu_max_11
 G cdu_min_11
 L cdu_sulfur_11
 L tank_farm_11
 L berth_12
 E tank_balance_arab_light_12
 E tank_balance_arab_heavy_12
 E tank_balance_basrah_medium_12
 E tank_balance_murban_12
 E tank_balance_upper_zakum_12
 E tank_balance_kuwait_export_12
 L cdu_max_12
 G cdu_min_12
 L cdu_sulfur_12
 L tank_farm_12
 L berth_13
 E tank_balance_arab_light_13
 E tank_balance_arab_heavy_13
 E tank_balance_basrah_medium_13
 E tank_balance_murban_13
 E tank_balance_upper_zakum_13
 E tank_balance_kuwait_export_13
 L cdu_max_13
 G cdu_min_13
 L cdu_sulfur_13
 L tank_farm_13
 G cdu_stops
 G closing_stock
COLUMNS
 M170 'MARKER' 'INTORG'
 arrive_0_arab_light_0 lift_0 1
 arrive_0_arab_light_0 berth_0 1
 arrive_0_arab_light_0 tank_balance_arab_light_0 130
 arrive_0_arab_light_1 lift_0 1
 arrive_0_arab_light_1 berth_1 1
 arrive_0_arab_light_1 tank_balance_arab_light_1 130
 arrive_0_arab_light_2 obj -32.5
 arrive_0_arab_light_2 lift_0 1
 arrive_0_arab_light_2 berth_2 1
 arrive_0_arab_light_2 tank_balance_arab_light_2 130
 arrive_0_arab_light_3 obj -65
 arrive_0_arab_light_3 lift_0 1
 arrive_0_arab_light_3 berth_3 1
 arrive_0_arab_light_3 tank_balance_arab_light_3 130
 skip_0_arab_light obj -5000
 skip_0_arab_light lift_0 1
 arrive_1_arab_heavy_10 lift_1 1
 arrive_1_arab_heavy_10 berth_10 1
 arrive_1_arab_heavy_10 tank_balance_arab_heavy_10 130
 arrive_1_arab_heavy_11 lift_1 1
 arrive_1_arab_heavy_11 berth_11 1
 arrive_1_arab_heavy_11 tank_balance_arab_heavy_11 130
 arrive_1_arab_heavy_12 obj -32.5
 arrive_1_arab_heavy_12 lift_1 1
 arrive_1_arab_heavy_12 berth_12 1
 arrive_1_arab_heavy_12 tank_balance_arab_heavy_12 130
 arrive_1_arab_heavy_13 obj -65
 arrive_1_arab_heavy_13 lift_1 1
 arrive_1_arab_heavy_13 berth_13 1
 arrive_1_arab_heavy_13 tank_balance_arab_heavy_13 130
 skip_1_arab_heavy obj -5000
 skip_1_arab_heavy lift_1 1
 arrive_2_basrah_medium_8 lift_2 1
 arrive_2_basrah_medium_8 berth_8 1
 arrive_2_basrah_medium_8 tank_balance_basrah_medium_8 100
 arrive_2_basrah_medium_9 lift_2 1
 arrive_2_basrah_medium_9 berth_9 1
 arrive_2_basrah_medium_9 tank_balance_basrah_medium_9 100
 arrive_2_basrah_medium_10 obj -25
 arrive_2_basrah_medium_10 lift_2 1
 arrive_2_basrah_medium_10 berth_10 1
 arrive_2_basrah_medium_10 tank_balance_basrah_medium_10 100
 skip_2_basrah_medium obj -5000
 skip_2_basrah_medium lift_2 1
 arrive_3_murban_7 lift_3 1
 arrive_3_murban_7 berth_7 1
 arrive_3_murban_7 tank_balance_murban_7 130
 arrive_3_murban_8 lift_3 1
 arrive_3_murban_8 berth_8 1
 arrive_3_murban_8 tank_balance_murban_8 130
 arrive_3_murban_9 obj -32.5
 arrive_3_murban_9 lift_3 1
 arrive_3_murban_9 berth_9 1
 arrive_3_murban_9 tank_balance_murban_9 130
 arrive_3_murban_10 obj -65
 arrive_3_murban_10 lift_3 1
 arrive_3_murban_10 berth_10 1
 arrive_3_murban_10 tank_balance_murban_10 130
 arrive_3_murban_11 obj -97.5
 arrive_3_murban_11 lift_3 1
 arrive_3_murban_11 berth_11 1
 arrive_3_murban_11 tank_balance_murban_11 130
 skip_3_murban obj -5000
 skip_3_murban lift_3 1
 arrive_4_upper_zakum_10 lift_4 1
 arrive_4_upper_zakum_10 berth_10 1
 arrive_4_upper_zakum_10 tank_balance_upper_zakum_10 130
 arrive_4_upper_zakum_11 lift_4 1
 arrive_4_upper_zakum_11 berth_11 1
 arrive_4_upper_zakum_11 tank_balance_upper_zakum_11 130
 arrive_4_upper_zakum_12 obj -32.5
 arrive_4_upper_zakum_12 lift_4 1
 arrive_4_upper_zakum_12 berth_12 1
 arrive_4_upper_zakum_12 tank_balance_upper_zakum_12 130
 arrive_4_upper_zakum_13 obj -65
 arrive_4_upper_zakum_13 lift_4 1
 arrive_4_upper_zakum_13 berth_13 1
 arrive_4_upper_zakum_13 tank_balance_upper_zakum_13 130
 skip_4_upper_zakum obj -5000
 skip_4_upper_zakum lift_4 1
 arrive_5_kuwait_export_9 lift_5 1
 arrive_5_kuwait_export_9 berth_9 1
 arrive_5_kuwait_export_9 tank_balance_kuwait_export_9 130
 arrive_5_kuwait_export_10 lift_5 1
 arrive_5_kuwait_export_10 berth_10 1
 arrive_5_kuwait_export_10 tank_balance_kuwait_export_10 130
 arrive_5_kuwait_export_11 obj -32.5
 arrive_5_kuwait_export_11 lift_5 1
 arrive_5_kuwait_export_11 berth_11 1
 arrive_5_kuwait_export_11 tank_balance_kuwait_export_11 130
 arrive_5_kuwait_export_12 obj -65
 arrive_5_kuwait_export_12 lift_5 1
 arrive_5_kuwait_export_12 berth_12 1
 arrive_5_kuwait_export_12 tank_balance_kuwait_export_12 130
 skip_5_kuwait_export obj -5000
 skip_5_kuwait_export lift_5 1
 arrive_6_arab_light_4 lift_6 1
 arrive_6_arab_light_4 berth_4 1
 arrive_6_arab_light_4 tank_balance_arab_light_4 80
 arrive_6_arab_light_5 lift_6 1
 arrive_6_arab_light_5 berth_5 1
 arrive_6_arab_light_5 tank_balance_arab_light_5 80
 arrive_6_arab_light_6 obj -20
 arrive_6_arab_light_6 lift_6 1
 arrive_6_arab_light_6 berth_6 1
 arrive_6_arab_light_6 tank_balance_arab_light_6 80
 arrive_6_arab_light_7 obj -40
 arrive_6_arab_light_7 lift_6 1
 arrive_6_arab_light_7 berth_7 1
 arrive_6_arab_light_7 tank_balance_arab_light_7 80
 arrive_6_arab_light_8 obj -60
 arrive_6_arab_light_8 lift_6 1
 arrive_6_arab_light_8 berth_8 1
 arrive_6_arab_light_8 tank_balance_arab_light_8 80
 skip_6_arab_light obj -5000
 skip_6_arab_light lift_6 1
 arrive_7_arab_heavy_1 lift_7 1
 arrive_7_arab_heavy_1 berth_1 1
 arrive_7_arab_heavy_1 tank_balance_arab_heavy_1 100
 arrive_7_arab_heavy_2 lift_7 1
 arrive_7_arab_heavy_2 berth_2 1
 arrive_7_arab_heavy_2 tank_balance_arab_heavy_2 100
 arrive_7_arab_heavy_3 obj -25
 arrive_7_arab_heavy_3 lift_7 1
 arrive_7_arab_heavy_3 berth_3 1
 arrive_7_arab_heavy_3 tank_balance_arab_heavy_3 100
 arrive_7_arab_heavy_4 obj -50
 arrive_7_arab_heavy_4 lift_7 1
 arrive_7_arab_heavy_4 berth_4 1
 arrive_7_arab_heavy_4 tank_balance_arab_heavy_4 100
 skip_7_arab_heavy obj -5000
 skip_7_arab_heavy lift_7 1
 spot_0_arab_heavy_0 obj -270.475707392
 spot_0_arab_heavy_0 berth_0 1
 spot_0_arab_heavy_0 tank_balance_arab_heavy_0 90
 spot_0_arab_heavy_1 obj -195.737585157
 spot_0_arab_heavy_1 berth_1 1
 spot_0_arab_heavy_1 tank_balance_arab_heavy_1 90
 spot_0_arab_heavy_2 obj -150.033167113
 spot_0_arab_heavy_2 berth_2 1
 spot_0_arab_heavy_2 tank_balance_arab_heavy_2 90
 spot_0_arab_heavy_3 obj -183.547374782
 spot_0_arab_heavy_3 berth_3 1
 spot_0_arab_heavy_3 tank_balance_arab_heavy_3 90
 spot_0_arab_heavy_4 obj -260.654782374
 spot_0_arab_heavy_4 berth_4 1
 spot_0_arab_heavy_4 tank_balance_arab_heavy_4 90
 spot_0_arab_heavy_5 obj -212.576695939
 spot_0_arab_heavy_5 berth_5 1
 spot_0_arab_heavy_5 tank_balance_arab_heavy_5 90
 spot_0_arab_heavy_6 obj -299.074831795
 spot_0_arab_heavy_6 berth_6 1
 spot_0_arab_heavy_6 tank_balance_arab_heavy_6 90
 spot_0_arab_heavy_7 obj -216.604148874
 spot_0_arab_heavy_7 berth_7 1
 spot_0_arab_heavy_7 tank_balance_arab_heavy_7 90
 spot_0_arab_heavy_8 obj -209.914095667
 spot_0_arab_heavy_8 berth_8 1
 spot_0_arab_heavy_8 tank_balance_arab_heavy_8 90
 spot_0_arab_heavy_9 obj -224.331887516
 spot_0_arab_heavy_9 berth_9 1
 spot_0_arab_heavy_9 tank_balance_arab_heavy_9 90
 spot_0_arab_heavy_10 obj -234.949299892
 spot_0_arab_heavy_10 berth_10 1
 spot_0_arab_heavy_10 tank_balance_arab_heavy_10 90
 spot_0_arab_heavy_11 obj -289.761437111
 spot_0_arab_heavy_11 berth_11 1
 spot_0_arab_heavy_11 tank_balance_arab_heavy_11 90
 spot_0_arab_heavy_12 obj -156.371788485
 spot_0_arab_heavy_12 berth_12 1
 spot_0_arab_heavy_12 tank_balance_arab_heavy_12 90
 spot_0_arab_heavy_13 obj -216.403314698
 spot_0_arab_heavy_13 berth_13 1
 spot_0_arab_heavy_13 tank_balance_arab_heavy_13 90
 spot_1_kuwait_export_0 obj -223.99053092
 spot_1_kuwait_export_0 berth_0 1
 spot_1_kuwait_export_0 tank_balance_kuwait_export_0 90
 spot_1_kuwait_export_1 obj -248.322329466
 spot_1_kuwait_export_1 berth_1 1
 spot_1_kuwait_export_1 tank_balance_kuwait_export_1 90
 spot_1_kuwait_export_2 obj -185.005616317
 spot_1_kuwait_export_2 berth_2 1
 spot_1_kuwait_export_2 tank_balance_kuwait_export_2 90
 spot_1_kuwait_export_3 obj -283.190116679
 spot_1_kuwait_export_3 berth_3 1
 spot_1_kuwait_export_3 tank_balance_kuwait_export_3 90
 spot_1_kuwait_export_4 obj -286.726910754
 spot_1_kuwait_export_4 berth_4 1
 spot_1_kuwait_export_4 tank_balance_kuwait_export_4 90
 spot_1_kuwait_export_5 obj -175.033616891
 spot_1_kuwait_export_5 berth_5 1
 spot_1_kuwait_export_5 tank_balance_kuwait_export_5 90
 spot_1_kuwait_export_6 obj -270.145834486
 spot_1_kuwait_export_6 berth_6 1
 spot_1_kuwait_export_6 tank_balance_kuwait_export_6 90
 spot_1_kuwait_export_7 obj -286.876596415
 spot_1_kuwait_export_7 berth_7 1
 spot_1_kuwait_export_7 tank_balance_kuwait_export_7 90
 spot_1_kuwait_export_8 obj -233.041248667
 spot_1_kuwait_export_8 berth_8 1
 spot_1_kuwait_export_8 tank_balance_kuwait_export_8 90
 spot_1_kuwait_export_9 obj -208.957532342
 spot_1_kuwait_export_9 berth_9 1
 spot_1_kuwait_export_9 tank_balance_kuwait_export_9 90
 spot_1_kuwait_export_10 obj -174.373917757
 spot_1_kuwait_export_10 berth_10 1
 spot_1_kuwait_export_10 tank_balance_kuwait_export_10 90
 spot_1_kuwait_export_11 obj -245.2824765
 spot_1_kuwait_export_11 berth_11 1
 spot_1_kuwait_export_11 tank_balance_kuwait_export_11 90
 spot_1_kuwait_export_12 obj -165.001414437
 spot_1_kuwait_export_12 berth_12 1
 spot_1_kuwait_export_12 tank_balance_kuwait_export_12 90
 spot_1_kuwait_export_13 obj -206.283374465
 spot_1_kuwait_export_13 berth_13 1
 spot_1_kuwait_export_13 tank_balance_kuwait_export_13 90
 M387 'MARKER' 'INTEND'
 tank_arab_light_0 obj -0.3
 tank_arab_light_0 tank_balance_arab_light_0 -1
 tank_arab_light_0 tank_farm_0 1
 tank_arab_light_0 tank_balance_arab_light_1 1
 charge_arab_light_0 obj 984
 charge_arab_light_0 tank_balance_arab_light_0 -1
 charge_arab_light_0 cdu_max_0 1
 charge_arab_light_0 cdu_min_0 1
 charge_arab_light_0 cdu_sulfur_0 -0.5
 tank_arab_heavy_0 obj -0.3
 tank_arab_heavy_0 tank_balance_arab_heavy_0 -1
 tank_arab_heavy_0 tank_farm_0 1
 tank_arab_heavy_0 tank_balance_arab_heavy_1 1
 charge_arab_heavy_0 obj 552
 charge_arab_heavy_0 tank_balance_arab_heavy_0 -1
 charge_arab_heavy_0 cdu_max_0 1
 charge_arab_heavy_0 cdu_min_0 1
 charge_arab_heavy_0 cdu_sulfur_0 0.6
 tank_basrah_medium_0 obj -0.3
 tank_basrah_medium_0 tank_balance_basrah_medium_0 -1
 tank_basrah_medium_0 tank_farm_0 1
 tank_basrah_medium_0 tank_balance_basrah_medium_1 1
 charge_basrah_medium_0 obj 612
 charge_basrah_medium_0 tank_balance_basrah_medium_0 -1
 charge_basrah_medium_0 cdu_max_0 1
 charge_basrah_medium_0 cdu_min_0 1
 charge_basrah_medium_0 cdu_sulfur_0 0.6
 tank_murban_0 obj -0.3
 tank_murban_0 tank_balance_murban_0 -1
 tank_murban_0 tank_farm_0 1
 tank_murban_0 tank_balance_murban_1 1
 charge_murban_0 obj 1524
 charge_murban_0 tank_balance_murban_0 -1
 charge_murban_0 cdu_max_0 1
 charge_murban_0 cdu_min_0 1
 charge_murban_0 cdu_sulfur_0 -1.5
 tank_upper_zakum_0 obj -0.3
 tank_upper_zakum_0 tank_balance_upper_zakum_0 -1
 tank_upper_zakum_0 tank_farm_0 1
 tank_upper_zakum_0 tank_balance_upper_zakum_1 1
 charge_upper_zakum_0 obj 1032
 charge_upper_zakum_0 tank_balance_upper_zakum_0 -1
 charge_upper_zakum_0 cdu_max_0 1
 charge_upper_zakum_0 cdu_min_0 1
 charge_upper_zakum_0 cdu_sulfur_0 -0.4
 tank_kuwait_export_0 obj -0.3
 tank_kuwait_export_0 tank_balance_kuwait_export_0 -1
 tank_kuwait_export_0 tank_farm_0 1
 tank_kuwait_export_0 tank_balance_kuwait_export_1 1
 charge_kuwait_export_0 obj 768
 charge_kuwait_export_0 tank_balance_kuwait_export_0 -1
 charge_kuwait_export_0 cdu_max_0 1
 charge_kuwait_export_0 cdu_min_0 1
 charge_kuwait_export_0 cdu_sulfur_0 0.3
 M442 'MARKER' 'INTORG'
 cdu_on_0 cdu_max_0 -45
 cdu_on_0 cdu_min_0 -36
 cdu_on_0 cdu_stops 1
 M446 'MARKER' 'INTEND'
 tank_arab_light_1 obj -0.3
 tank_arab_light_1 tank_balance_arab_light_1 -1
 tank_arab_light_1 tank_farm_1 1
 tank_arab_light_1 tank_balance_arab_light_2 1
 charge_arab_light_1 obj 984
 charge_arab_light_1 tank_balance_arab_light_1 -1
 charge_arab_light_1 cdu_max_1 1
 charge_arab_light_1 cdu_min_1 1
 charge_arab_light_1 cdu_sulfur_1 -0.5
 tank_arab_heavy_1 obj -0.3
 tank_arab_heavy_1 tank_balance_arab_heavy_1 -1
 tank_arab_heavy_1 tank_farm_1 1
 tank_arab_heavy_1 tank_balance_arab_heavy_2 1
 charge_arab_heavy_1 obj 552
 charge_arab_heavy_1 tank_balance_arab_heavy_1 -1
 charge_arab_heavy_1 cdu_max_1 1
 charge_arab_heavy_1 cdu_min_1 1
 charge_arab_heavy_1 cdu_sulfur_1 0.6
 tank_basrah_medium_1 obj -0.3
 tank_basrah_medium_1 tank_balance_basrah_medium_1 -1
 tank_basrah_medium_1 tank_farm_1 1
 tank_basrah_medium_1 tank_balance_basrah_medium_2 1
 charge_basrah_medium_1 obj 612
 charge_basrah_medium_1 tank_balance_basrah_medium_1 -1
 charge_basrah_medium_1 cdu_max_1 1
 charge_basrah_medium_1 cdu_min_1 1
 charge_basrah_medium_1 cdu_sulfur_1 0.6
 tank_murban_1 obj -0.3
 tank_murban_1 tank_balance_murban_1 -1
 tank_murban_1 tank_farm_1 1
 tank_murban_1 tank_balance_murban_2 1
 charge_murban_1 obj 1524
 charge_murban_1 tank_balance_murban_1 -1
 charge_murban_1 cdu_max_1 1
 charge_murban_1 cdu_min_1 1
 charge_murban_1 cdu_sulfur_1 -1.5
 tank_upper_zakum_1 obj -0.3
 tank_upper_zakum_1 tank_balance_upper_zakum_1 -1
 tank_upper_zakum_1 tank_farm_1 1
 tank_upper_zakum_1 tank_balance_upper_zakum_2 1
 charge_upper_zakum_1 obj 1032
 charge_upper_zakum_1 tank_balance_upper_zakum_1 -1
 charge_upper_zakum_1 cdu_max_1 1
 charge_upper_zakum_1 cdu_min_1 1
 charge_upper_zakum_1 cdu_sulfur_1 -0.4
 tank_kuwait_export_1 obj -0.3
 tank_kuwait_export_1 tank_balance_kuwait_export_1 -1
 tank_kuwait_export_1 tank_farm_1 1
 tank_kuwait_export_1 tank_balance_kuwait_export_2 1
 charge_kuwait_export_1 obj 768
 charge_kuwait_export_1 tank_balance_kuwait_export_1 -1
 charge_kuwait_export_1 cdu_max_1 1
 charge_kuwait_export_1 cdu_min_1 1
 charge_kuwait_export_1 cdu_sulfur_1 0.3
 M501 'MARKER' 'INTORG'
 cdu_on_1 cdu_max_1 -45
 cdu_on_1 cdu_min_1 -36
 cdu_on_1 cdu_stops 1
 M505 'MARKER' 'INTEND'
 tank_arab_light_2 obj -0.3
 tank_arab_light_2 tank_balance_arab_light_2 -1
 tank_arab_light_2 tank_farm_2 1
 tank_arab_light_2 tank_balance_arab_light_3 1
 charge_arab_light_2 obj 984
 charge_arab_light_2 tank_balance_arab_light_2 -1
 charge_arab_light_2 cdu_max_2 1
 charge_arab_light_2 cdu_min_2 1
 charge_arab_light_2 cdu_sulfur_2 -0.5
 tank_arab_heavy_2 obj -0.3
 tank_arab_heavy_2 tank_balance_arab_heavy_2 -1
 tank_arab_heavy_2 tank_farm_2 1
 tank_arab_heavy_2 tank_balance_arab_heavy_3 1
 charge_arab_heavy_2 obj 552
 charge_arab_heavy_2 tank_balance_arab_heavy_2 -1
 charge_arab_heavy_2 cdu_max_2 1
 charge_arab_heavy_2 cdu_min_2 1
 charge_arab_heavy_2 cdu_sulfur_2 0.6
 tank_basrah_medium_2 obj -0.3
 tank_basrah_medium_2 tank_balance_basrah_medium_2 -1
 tank_basrah_medium_2 tank_farm_2 1
 tank_basrah_medium_2 tank_balance_basrah_medium_3 1
 charge_basrah_medium_2 obj 612
 charge_basrah_medium_2 tank_balance_basrah_medium_2 -1
 charge_basrah_medium_2 cdu_max_2 1
 charge_basrah_medium_2 cdu_min_2 1
 charge_basrah_medium_2 cdu_sulfur_2 0.6
 tank_murban_2 obj -0.3
 tank_murban_2 tank_balance_murban_2 -1
 tank_murban_2 tank_farm_2 1
 tank_murban_2 tank_balance_murban_3 1
 charge_murban_2 obj 1524
 charge_murban_2 tank_balance_murban_2 -1
 charge_murban_2 cdu_max_2 1
 charge_murban_2 cdu_min_2 1
 charge_murban_2 cdu_sulfur_2 -1.5
 tank_upper_zakum_2 obj -0.3
 tank_upper_zakum_2 tank_balance_upper_zakum_2 -1
 tank_upper_zakum_2 tank_farm_2 1
 tank_upper_zakum_2 tank_balance_upper_zakum_3 1
 charge_upper_zakum_2 obj 1032
 charge_upper_zakum_2 tank_balance_upper_zakum_2 -1
 charge_upper_zakum_2 cdu_max_2 1
 charge_upper_zakum_2 cdu_min_2 1
 charge_upper_zakum_2 cdu_sulfur_2 -0.4
 tank_kuwait_export_2 obj -0.3
 tank_kuwait_export_2 tank_balance_kuwait_export_2 -1
 tank_kuwait_export_2 tank_farm_2 1
 tank_kuwait_export_2 tank_balance_kuwait_export_3 1
 charge_kuwait_export_2 obj 768
 charge_kuwait_export_2 tank_balance_kuwait_export_2 -1
 charge_kuwait_export_2 cdu_max_2 1
 charge_kuwait_export_2 cdu_min_2 1
 charge_kuwait_export_2 cdu_sulfur_2 0.3
 M560 'MARKER' 'INTORG'
 cdu_on_2 cdu_max_2 -45
 cdu_on_2 cdu_min_2 -36
 cdu_on_2 cdu_stops 1
 M564 'MARKER' 'INTEND'
 tank_arab_light_3 obj -0.3
 tank_arab_light_3 tank_balance_arab_light_3 -1
 tank_arab_light_3 tank_farm_3 1
 tank_arab_light_3 tank_balance_arab_light_4 1
 charge_arab_light_3 obj 984
 charge_arab_light_3 tank_balance_arab_light_3 -1
 charge_arab_light_3 cdu_max_3 1
 charge_arab_light_3 cdu_min_3 1
 charge_arab_light_3 cdu_sulfur_3 -0.5
 tank_arab_heavy_3 obj -0.3
 tank_arab_heavy_3 tank_balance_arab_heavy_3 -1
 tank_arab_heavy_3 tank_farm_3 1
 tank_arab_heavy_3 tank_balance_arab_heavy_4 1
 charge_arab_heavy_3 obj 552
 charge_arab_heavy_3 tank_balance_arab_heavy_3 -1
 charge_arab_heavy_3 cdu_max_3 1
 charge_arab_heavy_3 cdu_min_3 1
 charge_arab_heavy_3 cdu_sulfur_3 0.6
 tank_basrah_medium_3 obj -0.3
 tank_basrah_medium_3 tank_balance_basrah_medium_3 -1
 tank_basrah_medium_3 tank_farm_3 1
 tank_basrah_medium_3 tank_balance_basrah_medium_4 1
 charge_basrah_medium_3 obj 612
 charge_basrah_medium_3 tank_balance_basrah_medium_3 -1
 charge_basrah_medium_3 cdu_max_3 1
 charge_basrah_medium_3 cdu_min_3 1
 charge_basrah_medium_3 cdu_sulfur_3 0.6
 tank_murban_3 obj -0.3
 tank_murban_3 tank_balance_murban_3 -1
 tank_murban_3 tank_farm_3 1
 tank_murban_3 tank_balance_murban_4 1
 charge_murban_3 obj 1524
 charge_murban_3 tank_balance_murban_3 -1
 charge_murban_3 cdu_max_3 1
 charge_murban_3 cdu_min_3 1
 charge_murban_3 cdu_sulfur_3 -1.5
 tank_upper_zakum_3 obj -0.3
 tank_upper_zakum_3 tank_balance_upper_zakum_3 -1
 tank_upper_zakum_3 tank_farm_3 1
 tank_upper_zakum_3 tank_balance_upper_zakum_4 1
 charge_upper_zakum_3 obj 1032
 charge_upper_zakum_3 tank_balance_upper_zakum_3 -1
 charge_upper_zakum_3 cdu_max_3 1
 charge_upper_zakum_3 cdu_min_3 1
 charge_upper_zakum_3 cdu_sulfur_3 -0.4
 tank_kuwait_export_3 obj -0.3
 tank_kuwait_export_3 tank_balance_kuwait_export_3 -1
 tank_kuwait_export_3 tank_farm_3 1
 tank_kuwait_export_3 tank_balance_kuwait_export_4 1
 charge_kuwait_export_3 obj 768
 charge_kuwait_export_3 tank_balance_kuwait_export_3 -1
 charge_kuwait_export_3 cdu_max_3 1
 charge_kuwait_export_3 cdu_min_3 1
 charge_kuwait_export_3 cdu_sulfur_3 0.3
 M619 'MARKER' 'INTORG'
 cdu_on_3 cdu_max_3 -45
 cdu_on_3 cdu_min_3 -36
 cdu_on_3 cdu_stops 1
 M623 'MARKER' 'INTEND'
 tank_arab_light_4 obj -0.3
 tank_arab_light_4 tank_balance_arab_light_4 -1
 tank_arab_light_4 tank_farm_4 1
 tank_arab_light_4 tank_balance_arab_light_5 1
 charge_arab_light_4 obj 984
 charge_arab_light_4 tank_balance_arab_light_4 -1
 charge_arab_light_4 cdu_max_4 1
 charge_arab_light_4 cdu_min_4 1
 charge_arab_light_4 cdu_sulfur_4 -0.5
 tank_arab_heavy_4 obj -0.3
 tank_arab_heavy_4 tank_balance_arab_heavy_4 -1
 tank_arab_heavy_4 tank_farm_4 1
 tank_arab_heavy_4 tank_balance_arab_heavy_5 1
 charge_arab_heavy_4 obj 552
 charge_arab_heavy_4 tank_balance_arab_heavy_4 -1
 charge_arab_heavy_4 cdu_max_4 1
 charge_arab_heavy_4 cdu_min_4 1
 charge_arab_heavy_4 cdu_sulfur_4 0.6
 tank_basrah_medium_4 obj -0.3
 tank_basrah_medium_4 tank_balance_basrah_medium_4 -1
 tank_basrah_medium_4 tank_farm_4 1
 tank_basrah_medium_4 tank_balance_basrah_medium_5 1
 charge_basrah_medium_4 obj 612
 charge_basrah_medium_4 tank_balance_basrah_medium_4 -1
 charge_basrah_medium_4 cdu_max_4 1
 charge_basrah_medium_4 cdu_min_4 1
 charge_basrah_medium_4 cdu_sulfur_4 0.6
 tank_murban_4 obj -0.3
 tank_murban_4 tank_balance_murban_4 -1
 tank_murban_4 tank_farm_4 1
 tank_murban_4 tank_balance_murban_5 1
 charge_murban_4 obj 1524
 charge_murban_4 tank_balance_murban_4 -1
 charge_murban_4 cdu_max_4 1
 charge_murban_4 cdu_min_4 1
 charge_murban_4 cdu_sulfur_4 -1.5
 tank_upper_zakum_4 obj -0.3
 tank_upper_zakum_4 tank_balance_upper_zakum_4 -1
 tank_upper_zakum_4 tank_farm_4 1
 tank_upper_zakum_4 tank_balance_upper_zakum_5 1
 charge_upper_zakum_4 obj 1032
 charge_upper_zakum_4 tank_balance_upper_zakum_4 -1
 charge_upper_zakum_4 cdu_max_4 1
 charge_upper_zakum_4 cdu_min_4 1
 charge_upper_zakum_4 cdu_sulfur_4 -0.4
 tank_kuwait_export_4 obj -0.3
 tank_kuwait_export_4 tank_balance_kuwait_export_4 -1
 tank_kuwait_export_4 tank_farm_4 1
 tank_kuwait_export_4 tank_balance_kuwait_export_5 1
 charge_kuwait_export_4 obj 768
 charge_kuwait_export_4 tank_balance_kuwait_export_4 -1
 charge_kuwait_export_4 cdu_max_4 1
 charge_kuwait_export_4 cdu_min_4 1
 charge_kuwait_export_4 cdu_sulfur_4 0.3
 M678 'MARKER' 'INTORG'
 cdu_on_4 cdu_max_4 -45
 cdu_on_4 cdu_min_4 -36
 cdu_on_4 cdu_stops 1
 M682 'MARKER' 'INTEND'
 tank_arab_light_5 obj -0.3
 tank_arab_light_5 tank_balance_arab_light_5 -1
 tank_arab_light_5 tank_farm_5 1
 tank_arab_light_5 tank_balance_arab_light_6 1
 charge_arab_light_5 obj 984
 charge_arab_light_5 tank_balance_arab_light_5 -1
 charge_arab_light_5 cdu_max_5 1
 charge_arab_light_5 cdu_min_5 1
 charge_arab_light_5 cdu_sulfur_5 -0.5
 tank_arab_heavy_5 obj -0.3
 tank_arab_heavy_5 tank_balance_arab_heavy_5 -1
 tank_arab_heavy_5 tank_farm_5 1
 tank_arab_heavy_5 tank_balance_arab_heavy_6 1
 charge_arab_heavy_5 obj 552
 charge_arab_heavy_5 tank_balance_arab_heavy_5 -1
 charge_arab_heavy_5 cdu_max_5 1
 charge_arab_heavy_5 cdu_min_5 1
 charge_arab_heavy_5 cdu_sulfur_5 0.6
 tank_basrah_medium_5 obj -0.3
 tank_basrah_medium_5 tank_balance_basrah_medium_5 -1
 tank_basrah_medium_5 tank_farm_5 1
 tank_basrah_medium_5 tank_balance_basrah_medium_6 1
 charge_basrah_medium_5 obj 612
 charge_basrah_medium_5 tank_balance_basrah_medium_5 -1
 charge_basrah_medium_5 cdu_max_5 1
 charge_basrah_medium_5 cdu_min_5 1
 charge_basrah_medium_5 cdu_sulfur_5 0.6
 tank_murban_5 obj -0.3
 tank_murban_5 tank_balance_murban_5 -1
 tank_murban_5 tank_farm_5 1
 tank_murban_5 tank_balance_murban_6 1
 charge_murban_5 obj 1524
 charge_murban_5 tank_balance_murban_5 -1
 charge_murban_5 cdu_max_5 1
 charge_murban_5 cdu_min_5 1
 charge_murban_5 cdu_sulfur_5 -1.5
 tank_upper_zakum_5 obj -0.3
 tank_upper_zakum_5 tank_balance_upper_zakum_5 -1
 tank_upper_zakum_5 tank_farm_5 1
 tank_upper_zakum_5 tank_balance_upper_zakum_6 1
 charge_upper_zakum_5 obj 1032
 charge_upper_zakum_5 tank_balance_upper_zakum_5 -1
 charge_upper_zakum_5 cdu_max_5 1
 charge_upper_zakum_5 cdu_min_5 1
 charge_upper_zakum_5 cdu_sulfur_5 -0.4
 tank_kuwait_export_5 obj -0.3
 tank_kuwait_export_5 tank_balance_kuwait_export_5 -1
 tank_kuwait_export_5 tank_farm_5 1
 tank_kuwait_export_5 tank_balance_kuwait_export_6 1
 charge_kuwait_export_5 obj 768
 charge_kuwait_export_5 tank_balance_kuwait_export_5 -1
 charge_kuwait_export_5 cdu_max_5 1
 charge_kuwait_export_5 cdu_min_5 1
 charge_kuwait_export_5 cdu_sulfur_5 0.3
 M737 'MARKER' 'INTORG'
 cdu_on_5 cdu_max_5 -45
 cdu_on_5 cdu_min_5 -36
 cdu_on_5 cdu_stops 1
 M741 'MARKER' 'INTEND'
 tank_arab_light_6 obj -0.3
 tank_arab_light_6 tank_balance_arab_light_6 -1
 tank_arab_light_6 tank_farm_6 1
 tank_arab_light_6 tank_balance_arab_light_7 1
 charge_arab_light_6 obj 984
 charge_arab_light_6 tank_balance_arab_light_6 -1
 charge_arab_light_6 cdu_max_6 1
 charge_arab_light_6 cdu_min_6 1
 charge_arab_light_6 cdu_sulfur_6 -0.5
 tank_arab_heavy_6 obj -0.3
 tank_arab_heavy_6 tank_balance_arab_heavy_6 -1
 tank_arab_heavy_6 tank_farm_6 1
 tank_arab_heavy_6 tank_balance_arab_heavy_7 1
 charge_arab_heavy_6 obj 552
 charge_arab_heavy_6 tank_balance_arab_heavy_6 -1
 charge_arab_heavy_6 cdu_max_6 1
 charge_arab_heavy_6 cdu_min_6 1
 charge_arab_heavy_6 cdu_sulfur_6 0.6
 tank_basrah_medium_6 obj -0.3
 tank_basrah_medium_6 tank_balance_basrah_medium_6 -1
 tank_basrah_medium_6 tank_farm_6 1
 tank_basrah_medium_6 tank_balance_basrah_medium_7 1
 charge_basrah_medium_6 obj 612
 charge_basrah_medium_6 tank_balance_basrah_medium_6 -1
 charge_basrah_medium_6 cdu_max_6 1
 charge_basrah_medium_6 cdu_min_6 1
 charge_basrah_medium_6 cdu_sulfur_6 0.6
 tank_murban_6 obj -0.3
 tank_murban_6 tank_balance_murban_6 -1
 tank_murban_6 tank_farm_6 1
 tank_murban_6 tank_balance_murban_7 1
 charge_murban_6 obj 1524
 charge_murban_6 tank_balance_murban_6 -1
 charge_murban_6 cdu_max_6 1
 charge_murban_6 cdu_min_6 1
 charge_murban_6 cdu_sulfur_6 -1.5
 tank_upper_zakum_6 obj -0.3
 tank_upper_zakum_6 tank_balance_upper_zakum_6 -1
 tank_upper_zakum_6 tank_farm_6 1
 tank_upper_zakum_6 tank_balance_upper_zakum_7 1
 charge_upper_zakum_6 obj 1032
 charge_upper_zakum_6 tank_balance_upper_zakum_6 -1
 charge_upper_zakum_6 cdu_max_6 1
 charge_upper_zakum_6 cdu_min_6 1
 charge_upper_zakum_6 cdu_sulfur_6 -0.4
 tank_kuwait_export_6 obj -0.3
 tank_kuwait_export_6 tank_balance_kuwait_export_6 -1
 tank_kuwait_export_6 tank_farm_6 1
 tank_kuwait_export_6 tank_balance_kuwait_export_7 1
 charge_kuwait_export_6 obj 768
 charge_kuwait_export_6 tank_balance_kuwait_export_6 -1
 charge_kuwait_export_6 cdu_max_6 1
 charge_kuwait_export_6 cdu_min_6 1
 charge_kuwait_export_6 cdu_sulfur_6 0.3
 M796 'MARKER' 'INTORG'
 cdu_on_6 cdu_max_6 -45
 cdu_on_6 cdu_min_6 -36
 cdu_on_6 cdu_stops 1
 M800 'MARKER' 'INTEND'
 tank_arab_light_7 obj -0.3
 tank_arab_light_7 tank_balance_arab_light_7 -1
 tank_arab_light_7 tank_farm_7 1
 tank_arab_light_7 tank_balance_arab_light_8 1
 charge_arab_light_7 obj 984
 charge_arab_light_7 tank_balance_arab_light_7 -1
 charge_arab_light_7 cdu_max_7 1
 charge_arab_light_7 cdu_min_7 1
 charge_arab_light_7 cdu_sulfur_7 -0.5
 tank_arab_heavy_7 obj -0.3
 tank_arab_heavy_7 tank_balance_arab_heavy_7 -1
 tank_arab_heavy_7 tank_farm_7 1
 tank_arab_heavy_7 tank_balance_arab_heavy_8 1
 charge_arab_heavy_7 obj 552
 charge_arab_heavy_7 tank_balance_arab_heavy_7 -1
 charge_arab_heavy_7 cdu_max_7 1
 charge_arab_heavy_7 cdu_min_7 1
 charge_arab_heavy_7 cdu_sulfur_7 0.6
 tank_basrah_medium_7 obj -0.3
 tank_basrah_medium_7 tank_balance_basrah_medium_7 -1
 tank_basrah_medium_7 tank_farm_7 1
 tank_basrah_medium_7 tank_balance_basrah_medium_8 1
 charge_basrah_medium_7 obj 612
 charge_basrah_medium_7 tank_balance_basrah_medium_7 -1
 charge_basrah_medium_7 cdu_max_7 1
 charge_basrah_medium_7 cdu_min_7 1
 charge_basrah_medium_7 cdu_sulfur_7 0.6
 tank_murban_7 obj -0.3
 tank_murban_7 tank_balance_murban_7 -1
 tank_murban_7 tank_farm_7 1
 tank_murban_7 tank_balance_murban_8 1
 charge_murban_7 obj 1524
 charge_murban_7 tank_balance_murban_7 -1
 charge_murban_7 cdu_max_7 1
 charge_murban_7 cdu_min_7 1
 charge_murban_7 cdu_sulfur_7 -1.5
 tank_upper_zakum_7 obj -0.3
 tank_upper_zakum_7 tank_balance_upper_zakum_7 -1
 tank_upper_zakum_7 tank_farm_7 1
 tank_upper_zakum_7 tank_balance_upper_zakum_8 1
 charge_upper_zakum_7 obj 1032
 charge_upper_zakum_7 tank_balance_upper_zakum_7 -1
 charge_upper_zakum_7 cdu_max_7 1
 charge_upper_zakum_7 cdu_min_7 1
 charge_upper_zakum_7 cdu_sulfur_7 -0.4
 tank_kuwait_export_7 obj -0.3
 tank_kuwait_export_7 tank_balance_kuwait_export_7 -1
 tank_kuwait_export_7 tank_farm_7 1
 tank_kuwait_export_7 tank_balance_kuwait_export_8 1
 charge_kuwait_export_7 obj 768
 charge_kuwait_export_7 tank_balance_kuwait_export_7 -1
 charge_kuwait_export_7 cdu_max_7 1
 charge_kuwait_export_7 cdu_min_7 1
 charge_kuwait_export_7 cdu_sulfur_7 0.3
 M855 'MARKER' 'INTORG'
 cdu_on_7 cdu_max_7 -45
 cdu_on_7 cdu_min_7 -36
 cdu_on_7 cdu_stops 1
 M859 'MARKER' 'INTEND'
 tank_arab_light_8 obj -0.3
 tank_arab_light_8 tank_balance_arab_light_8 -1
 tank_arab_light_8 tank_farm_8 1
 tank_arab_light_8 tank_balance_arab_light_9 1
 charge_arab_light_8 obj 984
 charge_arab_light_8 tank_balance_arab_light_8 -1
 charge_arab_light_8 cdu_max_8 1
 charge_arab_light_8 cdu_min_8 1
 charge_arab_light_8 cdu_sulfur_8 -0.5
 tank_arab_heavy_8 obj -0.3
 tank_arab_heavy_8 tank_balance_arab_heavy_8 -1
 tank_arab_heavy_8 tank_farm_8 1
 tank_arab_heavy_8 tank_balance_arab_heavy_9 1
 charge_arab_heavy_8 obj 552
 charge_arab_heavy_8 tank_balance_arab_heavy_8 -1
 charge_arab_heavy_8 cdu_max_8 1
 charge_arab_heavy_8 cdu_min_8 1
 charge_arab_heavy_8 cdu_sulfur_8 0.6
 tank_basrah_medium_8 obj -0.3
 tank_basrah_medium_8 tank_balance_basrah_medium_8 -1
 tank_basrah_medium_8 tank_farm_8 1
 tank_basrah_medium_8 tank_balance_basrah_medium_9 1
 charge_basrah_medium_8 obj 612
 charge_basrah_medium_8 tank_balance_basrah_medium_8 -1
 charge_basrah_medium_8 cdu_max_8 1
 charge_basrah_medium_8 cdu_min_8 1
 charge_basrah_medium_8 cdu_sulfur_8 0.6
 tank_murban_8 obj -0.3
 tank_murban_8 tank_balance_murban_8 -1
 tank_murban_8 tank_farm_8 1
 tank_murban_8 tank_balance_murban_9 1
 charge_murban_8 obj 1524
 charge_murban_8 tank_balance_murban_8 -1
 charge_murban_8 cdu_max_8 1
 charge_murban_8 cdu_min_8 1
 charge_murban_8 cdu_sulfur_8 -1.5
 tank_upper_zakum_8 obj -0.3
 tank_upper_zakum_8 tank_balance_upper_zakum_8 -1
 tank_upper_zakum_8 tank_farm_8 1
 tank_upper_zakum_8 tank_balance_upper_zakum_9 1
 charge_upper_zakum_8 obj 1032
 charge_upper_zakum_8 tank_balance_upper_zakum_8 -1
 charge_upper_zakum_8 cdu_max_8 1
 charge_upper_zakum_8 cdu_min_8 1
 charge_upper_zakum_8 cdu_sulfur_8 -0.4
 tank_kuwait_export_8 obj -0.3
 tank_kuwait_export_8 tank_balance_kuwait_export_8 -1
 tank_kuwait_export_8 tank_farm_8 1
 tank_kuwait_export_8 tank_balance_kuwait_export_9 1
 charge_kuwait_export_8 obj 768
 charge_kuwait_export_8 tank_balance_kuwait_export_8 -1
 charge_kuwait_export_8 cdu_max_8 1
 charge_kuwait_export_8 cdu_min_8 1
 charge_kuwait_export_8 cdu_sulfur_8 0.3
 M914 'MARKER' 'INTORG'
 cdu_on_8 cdu_max_8 -45
 cdu_on_8 cdu_min_8 -36
 cdu_on_8 cdu_stops 1
 M918 'MARKER' 'INTEND'
 tank_arab_light_9 obj -0.3
 tank_arab_light_9 tank_balance_arab_light_9 -1
 tank_arab_light_9 tank_farm_9 1
 tank_arab_light_9 tank_balance_arab_light_10 1
 charge_arab_light_9 obj 984
 charge_arab_light_9 tank_balance_arab_light_9 -1
 charge_arab_light_9 cdu_max_9 1
 charge_arab_light_9 cdu_min_9 1
 charge_arab_light_9 cdu_sulfur_9 -0.5
 tank_arab_heavy_9 obj -0.3
 tank_arab_heavy_9 tank_balance_arab_heavy_9 -1
 tank_arab_heavy_9 tank_farm_9 1
 tank_arab_heavy_9 tank_balance_arab_heavy_10 1
 charge_arab_heavy_9 obj 552
 charge_arab_heavy_9 tank_balance_arab_heavy_9 -1
 charge_arab_heavy_9 cdu_max_9 1
 charge_arab_heavy_9 cdu_min_9 1
 charge_arab_heavy_9 cdu_sulfur_9 0.6
 tank_basrah_medium_9 obj -0.3
 tank_basrah_medium_9 tank_balance_basrah_medium_9 -1
 tank_basrah_medium_9 tank_farm_9 1
 tank_basrah_medium_9 tank_balance_basrah_medium_10 1
 charge_basrah_medium_9 obj 612
 charge_basrah_medium_9 tank_balance_basrah_medium_9 -1
 charge_basrah_medium_9 cdu_max_9 1
 charge_basrah_medium_9 cdu_min_9 1
 charge_basrah_medium_9 cdu_sulfur_9 0.6
 tank_murban_9 obj -0.3
 tank_murban_9 tank_balance_murban_9 -1
 tank_murban_9 tank_farm_9 1
 tank_murban_9 tank_balance_murban_10 1
 charge_murban_9 obj 1524
 charge_murban_9 tank_balance_murban_9 -1
 charge_murban_9 cdu_max_9 1
 charge_murban_9 cdu_min_9 1
 charge_murban_9 cdu_sulfur_9 -1.5
 tank_upper_zakum_9 obj -0.3
 tank_upper_zakum_9 tank_balance_upper_zakum_9 -1
 tank_upper_zakum_9 tank_farm_9 1
 tank_upper_zakum_9 tank_balance_upper_zakum_10 1
 charge_upper_zakum_9 obj 1032
 charge_upper_zakum_9 tank_balance_upper_zakum_9 -1
 charge_upper_zakum_9 cdu_max_9 1
 charge_upper_zakum_9 cdu_min_9 1
 charge_upper_zakum_9 cdu_sulfur_9 -0.4
 tank_kuwait_export_9 obj -0.3
 tank_kuwait_export_9 tank_balance_kuwait_export_9 -1
 tank_kuwait_export_9 tank_farm_9 1
 tank_kuwait_export_9 tank_balance_kuwait_export_10 1
 charge_kuwait_export_9 obj 768
 charge_kuwait_export_9 tank_balance_kuwait_export_9 -1
 charge_kuwait_export_9 cdu_max_9 1
 charge_kuwait_export_9 cdu_min_9 1
 charge_kuwait_export_9 cdu_sulfur_9 0.3
 M973 'MARKER' 'INTORG'
 cdu_on_9 cdu_max_9 -45
 cdu_on_9 cdu_min_9 -36
 cdu_on_9 cdu_stops 1
 M977 'MARKER' 'INTEND'
 tank_arab_light_10 obj -0.3
 tank_arab_light_10 tank_balance_arab_light_10 -1
 tank_arab_light_10 tank_farm_10 1
 tank_arab_light_10 tank_balance_arab_light_11 1
 charge_arab_light_10 obj 984
 charge_arab_light_10 tank_balance_arab_light_10 -1
 charge_arab_light_10 cdu_max_10 1
 charge_arab_light_10 cdu_min_10 1
 charge_arab_light_10 cdu_sulfur_10 -0.5
 tank_arab_heavy_10 obj -0.3
 tank_arab_heavy_10 tank_balance_arab_heavy_10 -1
 tank_arab_heavy_10 tank_farm_10 1
 tank_arab_heavy_10 tank_balance_arab_heavy_11 1
 charge_arab_heavy_10 obj 552
 charge_arab_heavy_10 tank_balance_arab_heavy_10 -1
 charge_arab_heavy_10 cdu_max_10 1
 charge_arab_heavy_10 cdu_min_10 1
 charge_arab_heavy_10 cdu_sulfur_10 0.6
 tank_basrah_medium_10 obj -0.3
 tank_basrah_medium_10 tank_balance_basrah_medium_10 -1
 tank_basrah_medium_10 tank_farm_10 1
 tank_basrah_medium_10 tank_balance_basrah_medium_11 1
 charge_basrah_medium_10 obj 612
 charge_basrah_medium_10 tank_balance_basrah_medium_10 -1
 charge_basrah_medium_10 cdu_max_10 1
 charge_basrah_medium_10 cdu_min_10 1
 charge_basrah_medium_10 cdu_sulfur_10 0.6
 tank_murban_10 obj -0.3
 tank_murban_10 tank_balance_murban_10 -1
 tank_murban_10 tank_farm_10 1
 tank_murban_10 tank_balance_murban_11 1
 charge_murban_10 obj 1524
 charge_murban_10 tank_balance_murban_10 -1
 charge_murban_10 cdu_max_10 1
 charge_murban_10 cdu_min_10 1
 charge_murban_10 cdu_sulfur_10 -1.5
 tank_upper_zakum_10 obj -0.3
 tank_upper_zakum_10 tank_balance_upper_zakum_10 -1
 tank_upper_zakum_10 tank_farm_10 1
 tank_upper_zakum_10 tank_balance_upper_zakum_11 1
 charge_upper_zakum_10 obj 1032
 charge_upper_zakum_10 tank_balance_upper_zakum_10 -1
 charge_upper_zakum_10 cdu_max_10 1
 charge_upper_zakum_10 cdu_min_10 1
 charge_upper_zakum_10 cdu_sulfur_10 -0.4
 tank_kuwait_export_10 obj -0.3
 tank_kuwait_export_10 tank_balance_kuwait_export_10 -1
 tank_kuwait_export_10 tank_farm_10 1
 tank_kuwait_export_10 tank_balance_kuwait_export_11 1
 charge_kuwait_export_10 obj 768
 charge_kuwait_export_10 tank_balance_kuwait_export_10 -1
 charge_kuwait_export_10 cdu_max_10 1
 charge_kuwait_export_10 cdu_min_10 1
 charge_kuwait_export_10 cdu_sulfur_10 0.3
 M1032 'MARKER' 'INTORG'
 cdu_on_10 cdu_max_10 -45
 cdu_on_10 cdu_min_10 -36
 cdu_on_10 cdu_stops 1
 M1036 'MARKER' 'INTEND'
 tank_arab_light_11 obj -0.3
 tank_arab_light_11 tank_balance_arab_light_11 -1
 tank_arab_light_11 tank_farm_11 1
 tank_arab_light_11 tank_balance_arab_light_12 1
 charge_arab_light_11 obj 984
 charge_arab_light_11 tank_balance_arab_light_11 -1
 charge_arab_light_11 cdu_max_11 1
 charge_arab_light_11 cdu_min_11 1
 charge_arab_light_11 cdu_sulfur_11 -0.5
 tank_arab_heavy_11 obj -0.3
 tank_arab_heavy_11 tank_balance_arab_heavy_11 -1
 tank_arab_heavy_11 tank_farm_11 1
 tank_arab_heavy_11 tank_balance_arab_heavy_12 1
 charge_arab_heavy_11 obj 552
 charge_arab_heavy_11 tank_balance_arab_heavy_11 -1
 charge_arab_heavy_11 cdu_max_11 1
 charge_arab_heavy_11 cdu_min_11 1
 charge_arab_heavy_11 cdu_sulfur_11 0.6
 tank_basrah_medium_11 obj -0.3
 tank_basrah_medium_11 tank_balance_basrah_medium_11 -1
 tank_basrah_medium_11 tank_farm_11 1
 tank_basrah_medium_11 tank_balance_basrah_medium_12 1
 charge_basrah_medium_11 obj 612
 charge_basrah_medium_11 tank_balance_basrah_medium_11 -1
 charge_basrah_medium_11 cdu_max_11 1
 charge_basrah_medium_11 cdu_min_11 1
 charge_basrah_medium_11 cdu_sulfur_11 0.6
 tank_murban_11 obj -0.3
 tank_murban_11 tank_balance_murban_11 -1
 tank_murban_11 tank_farm_11 1
 tank_murban_11 tank_balance_murban_12 1
 charge_murban_11 obj 1524
 charge_murban_11 tank_balance_murban_11 -1
 charge_murban_11 cdu_max_11 1
 charge_murban_11 cdu_min_11 1
 charge_murban_11 cdu_sulfur_11 -1.5
 tank_upper_zakum_11 obj -0.3
 tank_upper_zakum_11 tank_balance_upper_zakum_11 -1
 tank_upper_zakum_11 tank_farm_11 1
 tank_upper_zakum_11 tank_balance_upper_zakum_12 1
 charge_upper_zakum_11 obj 1032
 charge_upper_zakum_11 tank_balance_upper_zakum_11 -1
 charge_upper_zakum_11 cdu_max_11 1
 charge_upper_zakum_11 cdu_min_11 1
 charge_upper_zakum_11 cdu_sulfur_11 -0.4
 tank_kuwait_export_11 obj -0.3
 tank_kuwait_export_11 tank_balance_kuwait_export_11 -1
 tank_kuwait_export_11 tank_farm_11 1
 tank_kuwait_export_11 tank_balance_kuwait_export_12 1
 charge_kuwait_export_11 obj 768
 charge_kuwait_export_11 tank_balance_kuwait_export_11 -1
 charge_kuwait_export_11 cdu_max_11 1
 charge_kuwait_export_11 cdu_min_11 1
 charge_kuwait_export_11 cdu_sulfur_11 0.3
 M1091 'MARKER' 'INTORG'
 cdu_on_11 cdu_max_11 -45
 cdu_on_11 cdu_min_11 -36
 cdu_on_11 cdu_stops 1
 M1095 'MARKER' 'INTEND'
 tank_arab_light_12 obj -0.3
 tank_arab_light_12 tank_balance_arab_light_12 -1
 tank_arab_light_12 tank_farm_12 1
 tank_arab_light_12 tank_balance_arab_light_13 1
 charge_arab_light_12 obj 984
 charge_arab_light_12 tank_balance_arab_light_12 -1
 charge_arab_light_12 cdu_max_12 1
 charge_arab_light_12 cdu_min_12 1
 charge_arab_light_12 cdu_sulfur_12 -0.5
 tank_arab_heavy_12 obj -0.3
 tank_arab_heavy_12 tank_balance_arab_heavy_12 -1
 tank_arab_heavy_12 tank_farm_12 1
 tank_arab_heavy_12 tank_balance_arab_heavy_13 1
 charge_arab_heavy_12 obj 552
 charge_arab_heavy_12 tank_balance_arab_heavy_12 -1
 charge_arab_heavy_12 cdu_max_12 1
 charge_arab_heavy_12 cdu_min_12 1
 charge_arab_heavy_12 cdu_sulfur_12 0.6
 tank_basrah_medium_12 obj -0.3
 tank_basrah_medium_12 tank_balance_basrah_medium_12 -1
 tank_basrah_medium_12 tank_farm_12 1
 tank_basrah_medium_12 tank_balance_basrah_medium_13 1
 charge_basrah_medium_12 obj 612
 charge_basrah_medium_12 tank_balance_basrah_medium_12 -1
 charge_basrah_medium_12 cdu_max_12 1
 charge_basrah_medium_12 cdu_min_12 1
 charge_basrah_medium_12 cdu_sulfur_12 0.6
 tank_murban_12 obj -0.3
 tank_murban_12 tank_balance_murban_12 -1
 tank_murban_12 tank_farm_12 1
 tank_murban_12 tank_balance_murban_13 1
 charge_murban_12 obj 1524
 charge_murban_12 tank_balance_murban_12 -1
 charge_murban_12 cdu_max_12 1
 charge_murban_12 cdu_min_12 1
 charge_murban_12 cdu_sulfur_12 -1.5
 tank_upper_zakum_12 obj -0.3
 tank_upper_zakum_12 tank_balance_upper_zakum_12 -1
 tank_upper_zakum_12 tank_farm_12 1
 tank_upper_zakum_12 tank_balance_upper_zakum_13 1
 charge_upper_zakum_12 obj 1032
 charge_upper_zakum_12 tank_balance_upper_zakum_12 -1
 charge_upper_zakum_12 cdu_max_12 1
 charge_upper_zakum_12 cdu_min_12 1
 charge_upper_zakum_12 cdu_sulfur_12 -0.4
 tank_kuwait_export_12 obj -0.3
 tank_kuwait_export_12 tank_balance_kuwait_export_12 -1
 tank_kuwait_export_12 tank_farm_12 1
 tank_kuwait_export_12 tank_balance_kuwait_export_13 1
 charge_kuwait_export_12 obj 768
 charge_kuwait_export_12 tank_balance_kuwait_export_12 -1
 charge_kuwait_export_12 cdu_max_12 1
 charge_kuwait_export_12 cdu_min_12 1
 charge_kuwait_export_12 cdu_sulfur_12 0.3
 M1150 'MARKER' 'INTORG'
 cdu_on_12 cdu_max_12 -45
 cdu_on_12 cdu_min_12 -36
 cdu_on_12 cdu_stops 1
 M1154 'MARKER' 'INTEND'
 tank_arab_light_13 obj -0.3
 tank_arab_light_13 tank_balance_arab_light_13 -1
 tank_arab_light_13 tank_farm_13 1
 tank_arab_light_13 closing_stock 1
 charge_arab_light_13 obj 984
 charge_arab_light_13 tank_balance_arab_light_13 -1
 charge_arab_light_13 cdu_max_13 1
 charge_arab_light_13 cdu_min_13 1
 charge_arab_light_13 cdu_sulfur_13 -0.5
 tank_arab_heavy_13 obj -0.3
 tank_arab_heavy_13 tank_balance_arab_heavy_13 -1
 tank_arab_heavy_13 tank_farm_13 1
 tank_arab_heavy_13 closing_stock 1
 charge_arab_heavy_13 obj 552
 charge_arab_heavy_13 tank_balance_arab_heavy_13 -1
 charge_arab_heavy_13 cdu_max_13 1
 charge_arab_heavy_13 cdu_min_13 1
 charge_arab_heavy_13 cdu_sulfur_13 0.6
 tank_basrah_medium_13 obj -0.3
 tank_basrah_medium_13 tank_balance_basrah_medium_13 -1
 tank_basrah_medium_13 tank_farm_13 1
 tank_basrah_medium_13 closing_stock 1
 charge_basrah_medium_13 obj 612
 charge_basrah_medium_13 tank_balance_basrah_medium_13 -1
 charge_basrah_medium_13 cdu_max_13 1
 charge_basrah_medium_13 cdu_min_13 1
 charge_basrah_medium_13 cdu_sulfur_13 0.6
 tank_murban_13 obj -0.3
 tank_murban_13 tank_balance_murban_13 -1
 tank_murban_13 tank_farm_13 1
 tank_murban_13 closing_stock 1
 charge_murban_13 obj 1524
 charge_murban_13 tank_balance_murban_13 -1
 charge_murban_13 cdu_max_13 1
 charge_murban_13 cdu_min_13 1
 charge_murban_13 cdu_sulfur_13 -1.5
 tank_upper_zakum_13 obj -0.3
 tank_upper_zakum_13 tank_balance_upper_zakum_13 -1
 tank_upper_zakum_13 tank_farm_13 1
 tank_upper_zakum_13 closing_stock 1
 charge_upper_zakum_13 obj 1032
 charge_upper_zakum_13 tank_balance_upper_zakum_13 -1
 charge_upper_zakum_13 cdu_max_13 1
 charge_upper_zakum_13 cdu_min_13 1
 charge_upper_zakum_13 cdu_sulfur_13 -0.4
 tank_kuwait_export_13 obj -0.3
 tank_kuwait_export_13 tank_balance_kuwait_export_13 -1
 tank_kuwait_export_13 tank_farm_13 1
 tank_kuwait_export_13 closing_stock 1
 charge_kuwait_export_13 obj 768
 charge_kuwait_export_13 tank_balance_kuwait_export_13 -1
 charge_kuwait_export_13 cdu_max_13 1
 charge_kuwait_export_13 cdu_min_13 1
 charge_kuwait_export_13 cdu_sulfur_13 0.3
 M1209 'MARKER' 'INTORG'
 cdu_on_13 cdu_max_13 -45
 cdu_on_13 cdu_min_13 -36
 cdu_on_13 cdu_stops 1
 M1213 'MARKER' 'INTEND'
RHS
 rhs lift_0 1
 rhs lift_1 1
 rhs lift_2 1
 rhs lift_3 1
 rhs lift_4 1
 rhs lift_5 1
 rhs lift_6 1
 rhs lift_7 1
 rhs berth_0 1
 rhs tank_balance_arab_light_0 -42.9425039334
 rhs tank_balance_arab_heavy_0 -67.212352917
 rhs tank_balance_basrah_medium_0 -88.7895399298
 rhs tank_balance_murban_0 -44.3017125238
 rhs tank_balance_upper_zakum_0 -62.1687686322
 rhs tank_balance_kuwait_export_0 -63.1368100255
 rhs tank_farm_0 520
 rhs berth_1 1
 rhs tank_farm_1 520
 rhs berth_2 1
 rhs tank_farm_2 520
 rhs berth_3 1
 rhs tank_farm_3 520
 rhs berth_4 1
 rhs tank_farm_4 520
 rhs berth_5 1
 rhs tank_farm_5 520
 rhs berth_6 1
 rhs tank_farm_6 520
 rhs berth_7 1
 rhs tank_farm_7 520
 rhs berth_8 1
 rhs tank_farm_8 520
 rhs berth_9 1
 rhs tank_farm_9 520
 rhs berth_10 1
 rhs tank_farm_10 520
 rhs berth_11 1
 rhs tank_farm_11 520
 rhs berth_12 1
 rhs tank_farm_12 520
 rhs berth_13 1
 rhs tank_farm_13 520
 rhs cdu_stops 13
 rhs closing_stock 150
BOUNDS
 LO bnd arrive_0_arab_light_0 0
 UP bnd arrive_0_arab_light_0 1
 LO bnd arrive_0_arab_light_1 0
 UP bnd arrive_0_arab_light_1 1
 LO bnd arrive_0_arab_light_2 0
 UP bnd arrive_0_arab_light_2 1
 LO bnd arrive_0_arab_light_3 0
 UP bnd arrive_0_arab_light_3 1
 LO bnd skip_0_arab_light 0
 UP bnd skip_0_arab_light 1
 LO bnd arrive_1_arab_heavy_10 0
 UP bnd arrive_1_arab_heavy_10 1
 LO bnd arrive_1_arab_heavy_11 0
 UP bnd arrive_1_arab_heavy_11 1
 LO bnd arrive_1_arab_heavy_12 0
 UP bnd arrive_1_arab_heavy_12 1
 LO bnd arrive_1_arab_heavy_13 0
 UP bnd arrive_1_arab_heavy_13 1
 LO bnd skip_1_arab_heavy 0
 UP bnd skip_1_arab_heavy 1
 LO bnd arrive_2_basrah_medium_8 0
 UP bnd arrive_2_basrah_medium_8 1
 LO bnd arrive_2_basrah_medium_9 0
 UP bnd arrive_2_basrah_medium_9 1
 LO bnd arrive_2_basrah_medium_10 0
 UP bnd arrive_2_basrah_medium_10 1
 LO bnd skip_2_basrah_medium 0
 UP bnd skip_2_basrah_medium 1
 LO bnd arrive_3_murban_7 0
 UP bnd arrive_3_murban_7 1
 LO bnd arrive_3_murban_8 0
 UP bnd arrive_3_murban_8 1
 LO bnd arrive_3_murban_9 0
 UP bnd arrive_3_murban_9 1
 LO bnd arrive_3_murban_10 0
 UP bnd arrive_3_murban_10 1
 LO bnd arrive_3_murban_11 0
 UP bnd arrive_3_murban_11 1
 LO bnd skip_3_murban 0
 UP bnd skip_3_murban 1
 LO bnd arrive_4_upper_zakum_10 0
 UP bnd arrive_4_upper_zakum_10 1
 LO bnd arrive_4_upper_zakum_11 0
 UP bnd arrive_4_upper_zakum_11 1
 LO bnd arrive_4_upper_zakum_12 0
 UP bnd arrive_4_upper_zakum_12 1
 LO bnd arrive_4_upper_zakum_13 0
 UP bnd arrive_4_upper_zakum_13 1
 LO bnd skip_4_upper_zakum 0
 UP bnd skip_4_upper_zakum 1
 LO bnd arrive_5_kuwait_export_9 0
 UP bnd arrive_5_kuwait_export_9 1
 LO bnd arrive_5_kuwait_export_10 0
 UP bnd arrive_5_kuwait_export_10 1
 LO bnd arrive_5_kuwait_export_11 0
 UP bnd arrive_5_kuwait_export_11 1
 LO bnd arrive_5_kuwait_export_12 0
 UP bnd arrive_5_kuwait_export_12 1
 LO bnd skip_5_kuwait_export 0
 UP bnd skip_5_kuwait_export 1
 LO bnd arrive_6_arab_light_4 0
 UP bnd arrive_6_arab_light_4 1
 LO bnd arrive_6_arab_light_5 0
 UP bnd arrive_6_arab_light_5 1
 LO bnd arrive_6_arab_light_6 0
 UP bnd arrive_6_arab_light_6 1
 LO bnd arrive_6_arab_light_7 0
 UP bnd arrive_6_arab_light_7 1
 LO bnd arrive_6_arab_light_8 0
 UP bnd arrive_6_arab_light_8 1
 LO bnd skip_6_arab_light 0
 UP bnd skip_6_arab_light 1
 LO bnd arrive_7_arab_heavy_1 0
 UP bnd arrive_7_arab_heavy_1 1
 LO bnd arrive_7_arab_heavy_2 0
 UP bnd arrive_7_arab_heavy_2 1
 LO bnd arrive_7_arab_heavy_3 0
 UP bnd arrive_7_arab_heavy_3 1
 LO bnd arrive_7_arab_heavy_4 0
 UP bnd arrive_7_arab_heavy_4 1
 LO bnd skip_7_arab_heavy 0
 UP bnd skip_7_arab_heavy 1
 LO bnd spot_0_arab_heavy_0 0
 UP bnd spot_0_arab_heavy_0 1
 LO bnd spot_0_arab_heavy_1 0
 UP bnd spot_0_arab_heavy_1 1
 LO bnd spot_0_arab_heavy_2 0
 UP bnd spot_0_arab_heavy_2 1
 LO bnd spot_0_arab_heavy_3 0
 UP bnd spot_0_arab_heavy_3 1
 LO bnd spot_0_arab_heavy_4 0
 UP bnd spot_0_arab_heavy_4 1
 LO bnd spot_0_arab_heavy_5 0
 UP bnd spot_0_arab_heavy_5 1
 LO bnd spot_0_arab_heavy_6 0
 UP bnd spot_0_arab_heavy_6 1
 LO bnd spot_0_arab_heavy_7 0
 UP bnd spot_0_arab_heavy_7 1
 LO bnd spot_0_arab_heavy_8 0
 UP bnd spot_0_arab_heavy_8 1
 LO bnd spot_0_arab_heavy_9 0
 UP bnd spot_0_arab_heavy_9 1
 LO bnd spot_0_arab_heavy_10 0
 UP bnd spot_0_arab_heavy_10 1
 LO bnd spot_0_arab_heavy_11 0
 UP bnd spot_0_arab_heavy_11 1
 LO bnd spot_0_arab_heavy_12 0
 UP bnd spot_0_arab_heavy_12 1
 LO bnd spot_0_arab_heavy_13 0
 UP bnd spot_0_arab_heavy_13 1
 LO bnd spot_1_kuwait_export_0 0
 UP bnd spot_1_kuwait_export_0 1
 LO bnd spot_1_kuwait_export_1 0
 UP bnd spot_1_kuwait_export_1 1
 LO bnd spot_1_kuwait_export_2 0
 UP bnd spot_1_kuwait_export_2 1
 LO bnd spot_1_kuwait_export_3 0
 UP bnd spot_1_kuwait_export_3 1
 LO bnd spot_1_kuwait_export_4 0
 UP bnd spot_1_kuwait_export_4 1
 LO bnd spot_1_kuwait_export_5 0
 UP bnd spot_1_kuwait_export_5 1
 LO bnd spot_1_kuwait_export_6 0
 UP bnd spot_1_kuwait_export_6 1
 LO bnd spot_1_kuwait_export_7 0
 UP bnd spot_1_kuwait_export_7 1
 LO bnd spot_1_kuwait_export_8 0
 UP bnd spot_1_kuwait_export_8 1
 LO bnd spot_1_kuwait_export_9 0
 UP bnd spot_1_kuwait_export_9 1
 LO bnd spot_1_kuwait_export_10 0
 UP bnd spot_1_kuwait_export_10 1
 LO bnd spot_1_kuwait_export_11 0
 UP bnd spot_1_kuwait_export_11 1
 LO bnd spot_1_kuwait_export_12 0
 UP bnd spot_1_kuwait_export_12 1
 LO bnd spot_1_kuwait_export_13 0
 UP bnd spot_1_kuwait_export_13 1
 UP bnd tank_arab_light_0 180
 UP bnd tank_arab_heavy_0 180
 UP bnd tank_basrah_medium_0 180
 UP bnd tank_murban_0 180
 UP bnd tank_upper_zakum_0 180
 UP bnd tank_kuwait_export_0 180
 LO bnd cdu_on_0 0
 UP bnd cdu_on_0 1
 UP bnd tank_arab_light_1 180
 UP bnd tank_arab_heavy_1 180
 UP bnd tank_basrah_medium_1 180
 UP bnd tank_murban_1 180
 UP bnd tank_upper_zakum_1 180
 UP bnd tank_kuwait_export_1 180
 LO bnd cdu_on_1 0
 UP bnd cdu_on_1 1
 UP bnd tank_arab_light_2 180
 UP bnd tank_arab_heavy_2 180
 UP bnd tank_basrah_medium_2 180
 UP bnd tank_murban_2 180
 UP bnd tank_upper_zakum_2 180
 UP bnd tank_kuwait_export_2 180
 LO bnd cdu_on_2 0
 UP bnd cdu_on_2 1
 UP bnd tank_arab_light_3 180
 UP bnd tank_arab_heavy_3 180
 UP bnd tank_basrah_medium_3 180
 UP bnd tank_murban_3 180
 UP bnd tank_upper_zakum_3 180
 UP bnd tank_kuwait_export_3 180
 LO bnd cdu_on_3 0
 UP bnd cdu_on_3 1
 UP bnd tank_arab_light_4 180
 UP bnd tank_arab_heavy_4 180
 UP bnd tank_basrah_medium_4 180
 UP bnd tank_murban_4 180
 UP bnd tank_upper_zakum_4 180
 UP bnd tank_kuwait_export_4 180
 LO bnd cdu_on_4 0
 UP bnd cdu_on_4 1
 UP bnd tank_arab_light_5 180
 UP bnd tank_arab_heavy_5 180
 UP bnd tank_basrah_medium_5 180
 UP bnd tank_murban_5 180
 UP bnd tank_upper_zakum_5 180
 UP bnd tank_kuwait_export_5 180
 LO bnd cdu_on_5 0
 UP bnd cdu_on_5 1
 UP bnd tank_arab_light_6 180
 UP bnd tank_arab_heavy_6 180
 UP bnd tank_basrah_medium_6 180
 UP bnd tank_murban_6 180
 UP bnd tank_upper_zakum_6 180
 UP bnd tank_kuwait_export_6 180
 LO bnd cdu_on_6 0
 UP bnd cdu_on_6 1
 UP bnd tank_arab_light_7 180
 UP bnd tank_arab_heavy_7 180
 UP bnd tank_basrah_medium_7 180
 UP bnd tank_murban_7 180
 UP bnd tank_upper_zakum_7 180
 UP bnd tank_kuwait_export_7 180
 LO bnd cdu_on_7 0
 UP bnd cdu_on_7 1
 UP bnd tank_arab_light_8 180
 UP bnd tank_arab_heavy_8 180
 UP bnd tank_basrah_medium_8 180
 UP bnd tank_murban_8 180
 UP bnd tank_upper_zakum_8 180
 UP bnd tank_kuwait_export_8 180
 LO bnd cdu_on_8 0
 UP bnd cdu_on_8 1
 UP bnd tank_arab_light_9 180
 UP bnd tank_arab_heavy_9 180
 UP bnd tank_basrah_medium_9 180
 UP bnd tank_murban_9 180
 UP bnd tank_upper_zakum_9 180
 UP bnd tank_kuwait_export_9 180
 LO bnd cdu_on_9 0
 UP bnd cdu_on_9 1
 UP bnd tank_arab_light_10 180
 UP bnd tank_arab_heavy_10 180
 UP bnd tank_basrah_medium_10 180
 UP bnd tank_murban_10 180
 UP bnd tank_upper_zakum_10 180
 UP bnd tank_kuwait_export_10 180
 LO bnd cdu_on_10 0
 UP bnd cdu_on_10 1
 UP bnd tank_arab_light_11 180
 UP bnd tank_arab_heavy_11 180
 UP bnd tank_basrah_medium_11 180
 UP bnd tank_murban_11 180
 UP bnd tank_upper_zakum_11 180
 UP bnd tank_kuwait_export_11 180
 LO bnd cdu_on_11 0
 UP bnd cdu_on_11 1
 UP bnd tank_arab_light_12 180
 UP bnd tank_arab_heavy_12 180
 UP bnd tank_basrah_medium_12 180
 UP bnd tank_murban_12 180
 UP bnd tank_upper_zakum_12 180
 UP bnd tank_kuwait_export_12 180
 LO bnd cdu_on_12 0
 UP bnd cdu_on_12 1
 UP bnd tank_arab_light_13 180
 UP bnd tank_arab_heavy_13 180
 UP bnd tank_basrah_medium_13 180
 UP bnd tank_murban_13 180
 UP bnd tank_upper_zakum_13 180
 UP bnd tank_kuwait_export_13 180
 LO bnd cdu_on_13 0
 UP bnd cdu_on_13 1
ENDATA
