"""Enumerations and bit flags of the explicit-FEM path.

The numeric values are part of the interface (users OR the BC bits together and
compare domain/contact type codes), so they are kept identical to
flatworld/definitions.py:5-104 in the reference.
"""


class DomainType:
    """Domain type codes (flatworld/definitions.py:5-11)."""

    ANALYTICAL = 0b000001
    FEM = 0b000010
    RIGID = 0b000100
    SPRINGMASS = 0b001000
    HEIGHTFIELD = 0b100000
    VOXELMAP = 0b1000000


class ContactType:
    """Contact type = OR of the two domain types (flatworld/definitions.py:14-30).

    The misspellings are the reference's public names and are kept.
    """

    FLEXANLAYTICAL = 0b000011
    FLEXFLEX = 0b000010
    RIGIDANLAYTICAL = 0b000101
    FLEXRIGID = 0b000110
    RIGIDRIGID = 0b000100
    RIGIDSPRING = 0b001100
    FLEXSPRING = 0b001010
    ANALYTICALANALYTICAL = 0b000001
    ANALYTICALSPRING = 0b001001
    SPRINGSPRING = 0b001000
    FEXHEIGHTFIELD = 0b100010
    FLEXVOXELMAP = 0b1000010
    RIGIDHEIGHTFIELD = 0b100100
    RIGIDVOXELMAP = 0b1000100
    SPRINGHEIGHTFIELD = 0b101000
    SPRINGVOXELMAP = 0b1001000


class BoundaryConditionType:
    """Descriptive BC kinds (flatworld/definitions.py:33-47)."""

    FIXED = 1
    PRESCRIBED_DISPLACEMENT = 2
    PRESCRIBED_VELOCITY = 3
    PRESCRIBED_ACCELERATION = 4
    FORCE = 5
    GRAVITY = 6
    ENFORCE_DISPLACEMENT = 7
    ENFORCE_ROTATION_DISPLACEMENT = 8
    ENFORCE_VELOCITY = 9
    ENFORCE_ROTATION_VELOCITY = 10
    ENFORCE_ACCELERATION = 11
    ENFORCE_ROTATION_ACCELERATION = 12
    TORQUE = 13
    FIXED_ALL = 14


class RigidType:
    """Rigid primitive codes (flatworld/definitions.py:50-54)."""

    BALL = 0b00001
    BOX = 0b00010
    CAPSULE = 0b01000
    MESH = 0b10000


class MaterialType:
    """Constitutive law codes (flatworld/definitions.py:80-83)."""

    ELASTIC = 1
    NEOHOOKEAN = 2
    MISES = 3


# Collision filtering bits (flatworld/definitions.py:86-93).
COLLISION_CATEGORY_ROBOT = 0b00000001
COLLISION_CATEGORY_GROUND = 0b00000010
COLLISION_CATEGORY_FOOT = 0b00000100
COLLISION_CATEGORY_FEM = 0b00001000
COLLISION_CATEGORY_ORDINARY_RIGID = 0b00100000
COLLISION_CATEGORY_VIRTUAL = 0b00000000
COLLISION_MASK_ALL = 0b11111111

# Per-node boundary-condition bits (flatworld/definitions.py:95-104).
UTYPE = 0b000000001
VTYPE = 0b000000010
ATYPE = 0b000000100
FORCETYPE = 0b000001000
GRAVITY = 0b000010000
RTYPE = 0b000100000
ROTVTYPE = 0b001000000
ROTATYPE = 0b010000000
TORQUETYPE = 0b100000000

# Device-side node flag layout used by this package (no reference counterpart):
# bits 0..8 carry the BC bits above, bit 15 marks a mesh boundary node, and the
# upper 16 bits index the per-scene table of BC value vectors.
NODE_BC_MASK = 0x1FF
NODE_BOUNDARY_BIT = 1 << 15
NODE_BCV_SHIFT = 16
MAX_BC_VALUES = 1 << 16

__all__ = [
    "DomainType",
    "ContactType",
    "BoundaryConditionType",
    "RigidType",
    "MaterialType",
    "COLLISION_CATEGORY_ROBOT",
    "COLLISION_CATEGORY_GROUND",
    "COLLISION_CATEGORY_FOOT",
    "COLLISION_CATEGORY_FEM",
    "COLLISION_CATEGORY_ORDINARY_RIGID",
    "COLLISION_CATEGORY_VIRTUAL",
    "COLLISION_MASK_ALL",
    "UTYPE",
    "VTYPE",
    "ATYPE",
    "FORCETYPE",
    "GRAVITY",
    "RTYPE",
    "ROTVTYPE",
    "ROTATYPE",
    "TORQUETYPE",
]
