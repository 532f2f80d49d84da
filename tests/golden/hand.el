# hand-made: self loops, duplicates, reversed duplicates, isolated ids 1,4,6
0 2
2 0
0 2
3 3
2 5
5 7
7 7
0 7
9 5
% comment line
9 2
