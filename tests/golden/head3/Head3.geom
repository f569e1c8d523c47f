# Domain Description 1.1

Interfaces 3

Interface: "skull.3.tri"
Interface: "cortex.3.tri"
Interface: "scalp.3.tri"

Domains 4

Domain Scalp: 1 -3
Domain Brain: -2
Domain Air: 3
Domain Skull: 2 -1

