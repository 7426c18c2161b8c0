>A
A
>B
A
>C
G
>D
T
