// Reference fixture harness (test infrastructure, never shipped).
//
// Builds worlds from a scene spec (forge2d_amd/scenes.py) with the reference
// Box2D C API, steps them with b2World_Step and writes snapshots of the full
// world state after selected steps. The snapshot layout is the one
// forge2d_amd/snapshot.py parses and that the HIP engine exports/imports.
//
// It is compiled against the reference sources where they lie
// (oracle/Makefile) and links the reference objects, so it reads the
// reference's internal structures directly:
//   bodies/sims/states  S/body.h:13-159, S/world.h (b2World)
//   shapes              S/shape.h:13-52
//   contacts            S/contact.h:60-131
//   constraint graph    S/constraint_graph.h:22-45
//   broad-phase moves   S/broad_phase.h:74-82, contact id pool S/id_pool.h:8-12
//   broad-phase trees   I/collision.h:621-655 (b2DynamicTree), node record of
//                       S/dynamic_tree.c:19-50 (read through TreeNodeRec below)
//
// Modes:
//   ref_harness dump  <scene> <out> <world> <steps> <k1,k2,...>
//   ref_harness time  <scene> <steps> <warmup> <threads> [first_world count]
//   ref_harness pairs <scene> <out> <world> <steps>   (per-step created pair lists)
//   ref_harness geom  <scene> <out> <world>           (shape geometry as built by the reference)
//   ref_harness events <scene> <out> <world> <steps> <k1,k2,...>  (contact and body events after steps)
//   ref_harness query <scene> <queries> <out> <world> <steps>  (world queries after steps)
//   ref_harness qtime <scene> <queries> <steps> <threads> <first_world> <count>  (closest-ray timing)

#include "box2d/box2d.h"
#include "body.h"
#include "constraint_graph.h"
#include "contact.h"
#include "island.h"
#include "joint.h"
#include "shape.h"
#include "solver_set.h"
#include "world.h"

#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#pragma pack( push, 1 )
typedef struct
{
	int32_t body_count, shape_count, joint_count, flags;
	float gx, gy, restitution_threshold, hit_event_threshold, contact_hertz, contact_damping_ratio,
		max_contact_push_speed, maximum_linear_speed;
} SpecWorld;

typedef struct
{
	int32_t type;
	float px, py, qc, qs, vx, vy, w, lin_damp, ang_damp, gravity_scale, sleep_threshold;
	int32_t flags;
} SpecBody;

typedef struct
{
	int32_t body, kind;
	float friction, restitution, rolling, tangent_speed, density;
	uint32_t cat_lo, cat_hi, mask_lo, mask_hi;
	int32_t group, flags, npts;
	float radius;
	float data[16];
} SpecShape;

typedef struct
{
	int32_t body_a, body_b;
	float lax, lay, lbx, lby, ref_angle;
	int32_t enable_limit;
	float lower, upper;
	int32_t enable_motor;
	float motor_speed, max_motor_torque;
	int32_t enable_spring;
	float hertz, damping;
	int32_t collide_connected;
	float draw_size;
} SpecJoint;

// ---- snapshot records (forge2d_amd/snapshot.py mirrors these) ----
typedef struct
{
	int32_t magic, version, step, body_count, shape_count, contact_cap, joint_count, overflow_count, move_count,
		free_id_count, color_total, reserved0;
	float contact_speed, inv_h, reserved1, reserved2;
} SnapHeader; // 64 bytes

typedef struct
{
	int32_t type, set_index, local_index, flags;
	float px, py, qc, qs, cx, cy, c0x, c0y, q0c, q0s, lcx, lcy, vx, vy, w, fx, fy, torque, inv_mass, inv_i, min_ext,
		max_ext, sleep_time, reserved;
} SnapBody; // 16 + 24*4 = 112 bytes

typedef struct
{
	float aabb[4];
	float fat[4];
	int32_t enlarged, proxy_key, body, reserved;
} SnapShape; // 48 bytes

typedef struct
{
	float px, py, ax, ay, bx, by, sep, ni, ti, tni, nv;
	int32_t id, persisted;
} SnapPoint; // 52 bytes

typedef struct
{
	int32_t id, shape_a, shape_b, flags, color, local_index, sim_flags, point_count;
	float friction, restitution, rolling, tangent_speed, nx, ny, rolling_impulse, reserved;
	SnapPoint pts[2];
} SnapContact; // 64 + 104 = 168 bytes

typedef struct
{
	int32_t id, color, local_index, body_a, body_b, enable_limit, enable_motor, enable_spring;
	float linear_impulse_x, linear_impulse_y, spring_impulse, motor_impulse, lower_impulse, upper_impulse, ref_angle,
		reserved;
} SnapJoint; // 64 bytes
#pragma pack( pop )

typedef struct
{
	SpecWorld hdr;
	SpecBody* bodies;
	SpecShape* shapes;
	SpecJoint* joints;
} SceneWorld;

typedef struct
{
	int count;
	SceneWorld* worlds;
	unsigned char* buf;
} Scene;

static Scene load_scene( const char* path )
{
	Scene sc = { 0 };
	FILE* f = fopen( path, "rb" );
	if ( !f )
	{
		fprintf( stderr, "cannot open %s\n", path );
		exit( 2 );
	}
	fseek( f, 0, SEEK_END );
	long n = ftell( f );
	fseek( f, 0, SEEK_SET );
	sc.buf = malloc( n );
	if ( fread( sc.buf, 1, n, f ) != (size_t)n )
	{
		fprintf( stderr, "short read\n" );
		exit( 2 );
	}
	fclose( f );
	if ( memcmp( sc.buf, "F2DSCN01", 8 ) != 0 )
	{
		fprintf( stderr, "bad scene magic\n" );
		exit( 2 );
	}
	int32_t count;
	memcpy( &count, sc.buf + 8, 4 );
	sc.count = count;
	sc.worlds = calloc( count, sizeof( SceneWorld ) );
	long off = 16;
	for ( int i = 0; i < count; ++i )
	{
		SceneWorld* w = sc.worlds + i;
		memcpy( &w->hdr, sc.buf + off, sizeof( SpecWorld ) );
		off += sizeof( SpecWorld );
		w->bodies = (SpecBody*)( sc.buf + off );
		off += (long)w->hdr.body_count * sizeof( SpecBody );
		w->shapes = (SpecShape*)( sc.buf + off );
		off += (long)w->hdr.shape_count * sizeof( SpecShape );
		w->joints = (SpecJoint*)( sc.buf + off );
		off += (long)w->hdr.joint_count * sizeof( SpecJoint );
	}
	return sc;
}

static pthread_mutex_t g_world_lock = PTHREAD_MUTEX_INITIALIZER;

static b2ShapeId* g_shapeIds; // when set, build_world records the shapes' ids here
static b2JointId* g_jointIds; // and the joints' ids here

static b2WorldId build_world( const SceneWorld* sw, b2BodyId* bodyIds )
{
	b2WorldDef def = b2DefaultWorldDef();
	def.gravity = ( b2Vec2 ){ sw->hdr.gx, sw->hdr.gy };
	def.restitutionThreshold = sw->hdr.restitution_threshold;
	def.hitEventThreshold = sw->hdr.hit_event_threshold;
	def.contactHertz = sw->hdr.contact_hertz;
	def.contactDampingRatio = sw->hdr.contact_damping_ratio;
	def.maxContactPushSpeed = sw->hdr.max_contact_push_speed;
	def.maximumLinearSpeed = sw->hdr.maximum_linear_speed;
	def.enableSleep = ( sw->hdr.flags & 1 ) != 0;
	def.enableContinuous = ( sw->hdr.flags & 2 ) != 0;
	pthread_mutex_lock( &g_world_lock );
	b2WorldId wid = b2CreateWorld( &def );
	pthread_mutex_unlock( &g_world_lock );
	b2World_EnableWarmStarting( wid, ( sw->hdr.flags & 4 ) != 0 );

	for ( int i = 0; i < sw->hdr.body_count; ++i )
	{
		const SpecBody* sb = sw->bodies + i;
		b2BodyDef bd = b2DefaultBodyDef();
		bd.type = (b2BodyType)sb->type;
		bd.position = ( b2Vec2 ){ sb->px, sb->py };
		bd.rotation = ( b2Rot ){ sb->qc, sb->qs };
		bd.linearVelocity = ( b2Vec2 ){ sb->vx, sb->vy };
		bd.angularVelocity = sb->w;
		bd.linearDamping = sb->lin_damp;
		bd.angularDamping = sb->ang_damp;
		bd.gravityScale = sb->gravity_scale;
		bd.sleepThreshold = sb->sleep_threshold;
		bd.isBullet = ( sb->flags & 1 ) != 0;
		bd.fixedRotation = ( sb->flags & 2 ) != 0;
		bd.allowFastRotation = ( sb->flags & 4 ) != 0;
		bd.enableSleep = ( sb->flags & 8 ) != 0;
		bd.isAwake = ( sb->flags & 16 ) != 0;
		bodyIds[i] = b2CreateBody( wid, &bd );
	}
	for ( int i = 0; i < sw->hdr.shape_count; ++i )
	{
		const SpecShape* ss = sw->shapes + i;
		b2ShapeDef sd = b2DefaultShapeDef();
		sd.material.friction = ss->friction;
		sd.material.restitution = ss->restitution;
		sd.material.rollingResistance = ss->rolling;
		sd.material.tangentSpeed = ss->tangent_speed;
		sd.density = ss->density;
		sd.filter.categoryBits = (uint64_t)ss->cat_lo | ( (uint64_t)ss->cat_hi << 32 );
		sd.filter.maskBits = (uint64_t)ss->mask_lo | ( (uint64_t)ss->mask_hi << 32 );
		sd.filter.groupIndex = ss->group;
		sd.enableContactEvents = ( ss->flags & 1 ) != 0;
		sd.enableHitEvents = ( ss->flags & 2 ) != 0;
		sd.invokeContactCreation = ( ss->flags & 4 ) != 0;
		sd.enablePreSolveEvents = ( ss->flags & 8 ) != 0;
		sd.updateBodyMass = ( ss->flags & 16 ) != 0;
		b2BodyId body = bodyIds[ss->body];
		const float* d = ss->data;
		switch ( ss->kind )
		{
			case 0:
			{
				b2Circle c = { { d[0], d[1] }, ss->radius };
				b2ShapeId sid = b2CreateCircleShape( body, &sd, &c );
				if ( g_shapeIds ) g_shapeIds[i] = sid;
			}
			break;
			case 1:
			{
				b2Capsule c = { { d[0], d[1] }, { d[2], d[3] }, ss->radius };
				b2ShapeId sid = b2CreateCapsuleShape( body, &sd, &c );
				if ( g_shapeIds ) g_shapeIds[i] = sid;
			}
			break;
			case 2:
			{
				b2Segment s = { { d[0], d[1] }, { d[2], d[3] } };
				b2ShapeId sid = b2CreateSegmentShape( body, &sd, &s );
				if ( g_shapeIds ) g_shapeIds[i] = sid;
			}
			break;
			case 3:
			{
				b2Hull hull = b2ComputeHull( (const b2Vec2*)d, ss->npts );
				if ( hull.count == 0 )
				{
					fprintf( stderr, "degenerate hull in shape %d\n", i );
					exit( 3 );
				}
				b2Polygon p = b2MakePolygon( &hull, ss->radius );
				b2ShapeId sid = b2CreatePolygonShape( body, &sd, &p );
				if ( g_shapeIds ) g_shapeIds[i] = sid;
			}
			break;
			case 4:
			{
				b2Polygon p = b2MakeOffsetRoundedBox( d[0], d[1], ( b2Vec2 ){ d[2], d[3] }, ( b2Rot ){ d[4], d[5] },
													  ss->radius );
				b2ShapeId sid = b2CreatePolygonShape( body, &sd, &p );
				if ( g_shapeIds ) g_shapeIds[i] = sid;
			}
			break;
			default:
				fprintf( stderr, "bad shape kind\n" );
				exit( 3 );
		}
	}
	for ( int i = 0; i < sw->hdr.joint_count; ++i )
	{
		const SpecJoint* sj = sw->joints + i;
		b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
		jd.bodyIdA = bodyIds[sj->body_a];
		jd.bodyIdB = bodyIds[sj->body_b];
		jd.localAnchorA = ( b2Vec2 ){ sj->lax, sj->lay };
		jd.localAnchorB = ( b2Vec2 ){ sj->lbx, sj->lby };
		jd.referenceAngle = sj->ref_angle;
		jd.enableLimit = sj->enable_limit != 0;
		jd.lowerAngle = sj->lower;
		jd.upperAngle = sj->upper;
		jd.enableMotor = sj->enable_motor != 0;
		jd.motorSpeed = sj->motor_speed;
		jd.maxMotorTorque = sj->max_motor_torque;
		jd.enableSpring = sj->enable_spring != 0;
		jd.hertz = sj->hertz;
		jd.dampingRatio = sj->damping;
		jd.collideConnected = sj->collide_connected != 0;
		jd.drawSize = sj->draw_size;
		b2JointId jid = b2CreateRevoluteJoint( b2Body_GetWorld( jd.bodyIdA ), &jd );
		if ( g_jointIds ) g_jointIds[i] = jid;
	}
	return wid;
}

// Island and solver-set section of a version-2 snapshot (S/island.h:27-58,
// S/solver_set.h, S/id_pool.h): 16 header ints, then the island records, the
// island and solver-set free lists, the awake set's island and body orders, and
// per body the island links and contact-list head, per contact the island links,
// set index and both body-list edges, per joint the island links and set index.
static void write_islands( FILE* f, b2World* world )
{
	b2SolverSet* awake = world->solverSets.data + b2_awakeSet;
	int32_t hdr[16] = { 0 };
	hdr[0] = world->islands.count;
	hdr[1] = world->islandIdPool.freeArray.count;
	hdr[2] = awake->islandSims.count;
	hdr[3] = world->splitIslandId;
	hdr[4] = world->solverSets.count;
	hdr[5] = world->solverSetIdPool.freeArray.count;
	hdr[6] = awake->bodySims.count;
	hdr[7] = world->islandIdPool.nextIndex;
	hdr[8] = world->solverSetIdPool.nextIndex;
	fwrite( hdr, 4, 16, f );
	for ( int i = 0; i < world->islands.count; ++i )
	{
		const b2Island* is = world->islands.data + i;
		int32_t r[16] = { is->islandId,	   is->setIndex,	is->localIndex,	  is->headBody,
						  is->tailBody,	   is->bodyCount,	is->headContact,  is->tailContact,
						  is->contactCount, is->headJoint,	is->tailJoint,	  is->jointCount,
						  is->parentIsland, is->constraintRemoveCount, 0, 0 };
		fwrite( r, 4, 16, f );
	}
	fwrite( world->islandIdPool.freeArray.data, 4, world->islandIdPool.freeArray.count, f );
	for ( int i = 0; i < awake->islandSims.count; ++i )
		fwrite( &awake->islandSims.data[i].islandId, 4, 1, f );
	fwrite( world->solverSetIdPool.freeArray.data, 4, world->solverSetIdPool.freeArray.count, f );
	for ( int i = 0; i < awake->bodySims.count; ++i )
		fwrite( &awake->bodySims.data[i].bodyId, 4, 1, f );
	for ( int i = 0; i < world->bodies.count; ++i )
	{
		const b2Body* b = world->bodies.data + i;
		int32_t r[4] = { b->islandId, b->islandPrev, b->islandNext, b->headContactKey };
		fwrite( r, 4, 4, f );
	}
	for ( int i = 0; i < world->contacts.count; ++i )
	{
		const b2Contact* c = world->contacts.data + i;
		int32_t r[8] = { -1, -1, -1, -1, -1, -1, -1, -1 };
		if ( c->contactId != B2_NULL_INDEX )
		{
			r[0] = c->islandId;
			r[1] = c->islandPrev;
			r[2] = c->islandNext;
			r[3] = c->setIndex;
			r[4] = c->edges[0].prevKey;
			r[5] = c->edges[0].nextKey;
			r[6] = c->edges[1].prevKey;
			r[7] = c->edges[1].nextKey;
		}
		fwrite( r, 4, 8, f );
	}
	for ( int i = 0; i < world->joints.count; ++i )
	{
		const b2Joint* j = world->joints.data + i;
		int32_t r[4] = { -1, -1, -1, -1 };
		if ( j->jointId != B2_NULL_INDEX )
		{
			r[0] = j->islandId;
			r[1] = j->islandPrev;
			r[2] = j->islandNext;
			r[3] = j->setIndex;
		}
		fwrite( r, 4, 4, f );
	}
}

// The 40-byte node record of S/dynamic_tree.c:19-50 (its type is private to that file):
// box, category bits, children or user data, parent or free-list next, height, flags.
typedef struct
{
	float aabb[4];
	uint64_t categoryBits;
	int32_t child1, child2;
	int32_t parent;
	uint16_t height, flags;
} TreeNodeRec;
_Static_assert( sizeof( TreeNodeRec ) == 40, "tree node record" );

// Tree section of a version-3 snapshot: for the static, kinematic and dynamic trees
// (S/broad_phase.h:56-60) an int32 header [root, node count, proxy count, fresh index,
// freed count, 0, 0, 0], the node records [0, fresh index), then the freed node ids
// as a stack (bottom first, the next allocation last). The node free list of
// S/dynamic_tree.c:122-170 is those freed ids above the never-used ids [fresh index,
// capacity) in ascending order, so this pair describes it exactly.
static void write_trees( FILE* f, b2World* world )
{
	for ( int t = 0; t < b2_bodyTypeCount; ++t )
	{
		b2DynamicTree* tree = world->broadPhase.trees + t;
		const TreeNodeRec* nodes = (const TreeNodeRec*)tree->nodes;
		int cap = tree->nodeCapacity;
		int* list = malloc( sizeof( int ) * ( cap + 1 ) );
		int n = 0;
		for ( int k = tree->freeList; k != B2_NULL_INDEX && n < cap; k = nodes[k].parent )
		{
			list[n++] = k;
		}
		int fresh = cap, freed = n;
		if ( n > 0 && list[n - 1] == cap - 1 )
		{
			int j = n - 1;
			while ( j > 0 && list[j - 1] == list[j] - 1 )
			{
				j -= 1;
			}
			fresh = list[j];
			freed = j;
		}
		int32_t hdr[8] = { tree->root, tree->nodeCount, tree->proxyCount, fresh, freed, 0, 0, 0 };
		fwrite( hdr, 4, 8, f );
		for ( int i = 0; i < fresh; ++i )
		{
			TreeNodeRec r = nodes[i];
			if ( ( r.flags & b2_allocatedNode ) == 0 )
			{
				// a free node: only its place in the free list is state
				memset( &r, 0, sizeof( r ) );
				r.child1 = r.child2 = r.parent = B2_NULL_INDEX;
			}
			fwrite( &r, sizeof( r ), 1, f );
		}
		for ( int i = freed - 1; i >= 0; --i )
		{
			fwrite( list + i, 4, 1, f );
		}
		free( list );
	}
}

static void write_snapshot( FILE* f, b2WorldId wid, int step )
{
	b2World* world = b2GetWorldFromId( wid );
	SnapHeader h = { 0 };
	h.magic = 0x50414E53;
	h.version = 3;
	h.step = step;
	h.body_count = world->bodies.count;
	h.shape_count = world->shapes.count;
	h.contact_cap = world->contacts.count;
	h.joint_count = world->joints.count;
	h.overflow_count = world->constraintGraph.colors[B2_OVERFLOW_INDEX].contactSims.count;
	h.move_count = world->broadPhase.moveArray.count;
	h.free_id_count = world->contactIdPool.freeArray.count;
	int colorTotal = 0;
	for ( int c = 0; c < B2_GRAPH_COLOR_COUNT; ++c )
	{
		colorTotal += world->constraintGraph.colors[c].contactSims.count;
	}
	h.color_total = colorTotal;
	h.contact_speed = world->contactSpeed;
	h.inv_h = world->inv_h;
	fwrite( &h, sizeof( h ), 1, f );

	for ( int i = 0; i < world->bodies.count; ++i )
	{
		b2Body* body = world->bodies.data + i;
		SnapBody sb = { 0 };
		sb.type = body->type;
		sb.set_index = body->setIndex;
		sb.local_index = body->localIndex;
		b2BodySim* sim = b2GetBodySim( world, body );
		sb.flags = ( sim->isFast ? 1 : 0 ) | ( sim->isBullet ? 2 : 0 ) | ( body->isSpeedCapped ? 4 : 0 );
		sb.px = sim->transform.p.x;
		sb.py = sim->transform.p.y;
		sb.qc = sim->transform.q.c;
		sb.qs = sim->transform.q.s;
		sb.cx = sim->center.x;
		sb.cy = sim->center.y;
		sb.c0x = sim->center0.x;
		sb.c0y = sim->center0.y;
		sb.q0c = sim->rotation0.c;
		sb.q0s = sim->rotation0.s;
		sb.lcx = sim->localCenter.x;
		sb.lcy = sim->localCenter.y;
		b2BodyState* state = b2GetBodyState( world, body );
		if ( state != NULL )
		{
			sb.vx = state->linearVelocity.x;
			sb.vy = state->linearVelocity.y;
			sb.w = state->angularVelocity;
		}
		sb.fx = sim->force.x;
		sb.fy = sim->force.y;
		sb.torque = sim->torque;
		sb.inv_mass = sim->invMass;
		sb.inv_i = sim->invInertia;
		sb.min_ext = sim->minExtent;
		sb.max_ext = sim->maxExtent;
		sb.sleep_time = body->sleepTime;
		fwrite( &sb, sizeof( sb ), 1, f );
	}

	for ( int i = 0; i < world->shapes.count; ++i )
	{
		b2Shape* shape = world->shapes.data + i;
		SnapShape ss = { 0 };
		ss.aabb[0] = shape->aabb.lowerBound.x;
		ss.aabb[1] = shape->aabb.lowerBound.y;
		ss.aabb[2] = shape->aabb.upperBound.x;
		ss.aabb[3] = shape->aabb.upperBound.y;
		ss.fat[0] = shape->fatAABB.lowerBound.x;
		ss.fat[1] = shape->fatAABB.lowerBound.y;
		ss.fat[2] = shape->fatAABB.upperBound.x;
		ss.fat[3] = shape->fatAABB.upperBound.y;
		ss.enlarged = shape->enlargedAABB;
		ss.proxy_key = shape->proxyKey;
		ss.body = shape->bodyId;
		fwrite( &ss, sizeof( ss ), 1, f );
	}

	for ( int i = 0; i < world->contacts.count; ++i )
	{
		b2Contact* contact = world->contacts.data + i;
		SnapContact sc;
		memset( &sc, 0, sizeof( sc ) );
		sc.id = contact->contactId;
		if ( contact->contactId != B2_NULL_INDEX )
		{
			b2ContactSim* sim = b2GetContactSim( world, contact );
			sc.shape_a = contact->shapeIdA;
			sc.shape_b = contact->shapeIdB;
			sc.flags = (int32_t)contact->flags;
			sc.color = contact->colorIndex;
			sc.local_index = contact->localIndex;
			sc.sim_flags = (int32_t)sim->simFlags;
			sc.point_count = sim->manifold.pointCount;
			sc.friction = sim->friction;
			sc.restitution = sim->restitution;
			sc.rolling = sim->rollingResistance;
			sc.tangent_speed = sim->tangentSpeed;
			sc.nx = sim->manifold.normal.x;
			sc.ny = sim->manifold.normal.y;
			sc.rolling_impulse = sim->manifold.rollingImpulse;
			for ( int k = 0; k < 2; ++k )
			{
				const b2ManifoldPoint* mp = sim->manifold.points + k;
				SnapPoint* sp = sc.pts + k;
				sp->px = mp->point.x;
				sp->py = mp->point.y;
				sp->ax = mp->anchorA.x;
				sp->ay = mp->anchorA.y;
				sp->bx = mp->anchorB.x;
				sp->by = mp->anchorB.y;
				sp->sep = mp->separation;
				sp->ni = mp->normalImpulse;
				sp->ti = mp->tangentImpulse;
				sp->tni = mp->totalNormalImpulse;
				sp->nv = mp->normalVelocity;
				sp->id = mp->id;
				sp->persisted = mp->persisted;
			}
		}
		fwrite( &sc, sizeof( sc ), 1, f );
	}

	for ( int i = 0; i < world->joints.count; ++i )
	{
		b2Joint* joint = world->joints.data + i;
		SnapJoint sj;
		memset( &sj, 0, sizeof( sj ) );
		sj.id = joint->jointId;
		if ( joint->jointId != B2_NULL_INDEX )
		{
			b2JointSim* js = b2GetJointSim( world, joint );
			sj.color = joint->colorIndex;
			sj.local_index = joint->localIndex;
			sj.body_a = joint->edges[0].bodyId;
			sj.body_b = joint->edges[1].bodyId;
			if ( js->type == b2_prismaticJoint )
			{
				// kind 1 in the reserved slot; the block impulse in linear_impulse
				const b2PrismaticJoint* pj = &js->prismaticJoint;
				sj.enable_limit = pj->enableLimit;
				sj.enable_motor = pj->enableMotor;
				sj.enable_spring = pj->enableSpring;
				sj.linear_impulse_x = pj->impulse.x;
				sj.linear_impulse_y = pj->impulse.y;
				sj.spring_impulse = pj->springImpulse;
				sj.motor_impulse = pj->motorImpulse;
				sj.lower_impulse = pj->lowerImpulse;
				sj.upper_impulse = pj->upperImpulse;
				sj.ref_angle = pj->referenceAngle;
				sj.reserved = 1.0f;
			}
			else if ( js->type == b2_weldJoint )
			{
				// kind 2; the angular impulse in spring_impulse
				const b2WeldJoint* wj = &js->weldJoint;
				sj.linear_impulse_x = wj->linearImpulse.x;
				sj.linear_impulse_y = wj->linearImpulse.y;
				sj.spring_impulse = wj->angularImpulse;
				sj.ref_angle = wj->referenceAngle;
				sj.reserved = 2.0f;
			}
			else if ( js->type == b2_distanceJoint )
			{
				// kind 4; the length impulse in linear_impulse_x
				const b2DistanceJoint* dj = &js->distanceJoint;
				sj.enable_limit = dj->enableLimit;
				sj.enable_motor = dj->enableMotor;
				sj.enable_spring = dj->enableSpring;
				sj.linear_impulse_x = dj->impulse;
				sj.motor_impulse = dj->motorImpulse;
				sj.lower_impulse = dj->lowerImpulse;
				sj.upper_impulse = dj->upperImpulse;
				sj.reserved = 4.0f;
			}
			else if ( js->type == b2_wheelJoint )
			{
				// kind 3; the perpendicular impulse in linear_impulse_x
				const b2WheelJoint* hj = &js->wheelJoint;
				sj.enable_limit = hj->enableLimit;
				sj.enable_motor = hj->enableMotor;
				sj.enable_spring = hj->enableSpring;
				sj.linear_impulse_x = hj->perpImpulse;
				sj.spring_impulse = hj->springImpulse;
				sj.motor_impulse = hj->motorImpulse;
				sj.lower_impulse = hj->lowerImpulse;
				sj.upper_impulse = hj->upperImpulse;
				sj.reserved = 3.0f;
			}
			else if ( js->type == b2_motorJoint )
			{
				// kind 5; the angular impulse in spring_impulse, the angular offset in ref_angle
				const b2MotorJoint* mj = &js->motorJoint;
				sj.linear_impulse_x = mj->linearImpulse.x;
				sj.linear_impulse_y = mj->linearImpulse.y;
				sj.spring_impulse = mj->angularImpulse;
				sj.ref_angle = mj->angularOffset;
				sj.reserved = 5.0f;
			}
			else if ( js->type == b2_mouseJoint )
			{
				// kind 6; the angular impulse in spring_impulse
				const b2MouseJoint* mj = &js->mouseJoint;
				sj.linear_impulse_x = mj->linearImpulse.x;
				sj.linear_impulse_y = mj->linearImpulse.y;
				sj.spring_impulse = mj->angularImpulse;
				sj.reserved = 6.0f;
			}
			else if ( js->type == b2_filterJoint )
			{
				sj.reserved = 7.0f;
			}
			else
			{
				const b2RevoluteJoint* rj = &js->revoluteJoint;
				sj.enable_limit = rj->enableLimit;
				sj.enable_motor = rj->enableMotor;
				sj.enable_spring = rj->enableSpring;
				sj.linear_impulse_x = rj->linearImpulse.x;
				sj.linear_impulse_y = rj->linearImpulse.y;
				sj.spring_impulse = rj->springImpulse;
				sj.motor_impulse = rj->motorImpulse;
				sj.lower_impulse = rj->lowerImpulse;
				sj.upper_impulse = rj->upperImpulse;
				sj.ref_angle = rj->referenceAngle;
			}
		}
		fwrite( &sj, sizeof( sj ), 1, f );
	}

	// constraint graph colour arrays, in array order (overflow last)
	for ( int c = 0; c < B2_GRAPH_COLOR_COUNT; ++c )
	{
		int32_t n = world->constraintGraph.colors[c].contactSims.count;
		fwrite( &n, 4, 1, f );
	}
	for ( int c = 0; c < B2_GRAPH_COLOR_COUNT; ++c )
	{
		b2GraphColor* color = world->constraintGraph.colors + c;
		for ( int k = 0; k < color->contactSims.count; ++k )
		{
			int32_t id = color->contactSims.data[k].contactId;
			fwrite( &id, 4, 1, f );
		}
	}
	fwrite( world->broadPhase.moveArray.data, 4, world->broadPhase.moveArray.count, f );
	fwrite( world->contactIdPool.freeArray.data, 4, world->contactIdPool.freeArray.count, f );
	write_islands( f, world );
	write_trees( f, world );
}

static int parse_list( const char* s, int* out, int cap )
{
	int n = 0;
	while ( *s && n < cap )
	{
		out[n++] = atoi( s );
		while ( *s && *s != ',' )
			++s;
		if ( *s == ',' )
			++s;
	}
	return n;
}

static double now_sec( void )
{
	struct timespec ts;
	clock_gettime( CLOCK_MONOTONIC, &ts );
	return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// ---------------------------------------------------------------------------- timing

typedef struct
{
	const Scene* sc;
	int first, count, steps, warmup;
	int next;
	pthread_mutex_t lock;
	double* step_ms; // per world-step wall times (world-major)
	double body_substeps;
	int substeps;
} TimeJob;

static void* time_worker( void* arg )
{
	TimeJob* job = arg;
	for ( ;; )
	{
		pthread_mutex_lock( &job->lock );
		int k = job->next++;
		pthread_mutex_unlock( &job->lock );
		if ( k >= job->count )
			break;
		const SceneWorld* sw = job->sc->worlds + job->first + k;
		b2BodyId* ids = malloc( sizeof( b2BodyId ) * ( sw->hdr.body_count + 1 ) );
		b2WorldId wid = build_world( sw, ids );
		for ( int s = 0; s < job->warmup; ++s )
			b2World_Step( wid, 1.0f / 60.0f, job->substeps );
		for ( int s = 0; s < job->steps; ++s )
		{
			double t0 = now_sec();
			b2World_Step( wid, 1.0f / 60.0f, job->substeps );
			double t1 = now_sec();
			job->step_ms[(long)k * job->steps + s] = 1e3 * ( t1 - t0 );
		}
		b2Counters counters = b2World_GetCounters( wid );
		(void)counters;
		pthread_mutex_lock( &g_world_lock );
		b2DestroyWorld( wid );
		pthread_mutex_unlock( &g_world_lock );
		free( ids );
	}
	return NULL;
}

static int cmp_double( const void* a, const void* b )
{
	double x = *(const double*)a, y = *(const double*)b;
	return x < y ? -1 : ( x > y ? 1 : 0 );
}

// ---- world queries (forge2d_amd f2dQuery records, include/forge2d_amd.h) ----
typedef struct
{
	int32_t world, kind;
	uint32_t cat_lo, cat_hi, mask_lo, mask_hi;
	float p[6];
} QueryRec; // 48 bytes

typedef struct
{
	int32_t shape;
	float px, py, nx, ny, fraction;
} HitRec; // 24 bytes

typedef struct
{
	HitRec* hits;
	int count, cap;
} HitList;

static void hit_push( HitList* l, HitRec r )
{
	if ( l->count == l->cap )
	{
		l->cap = l->cap ? 2 * l->cap : 64;
		l->hits = realloc( l->hits, sizeof( HitRec ) * l->cap );
	}
	l->hits[l->count++] = r;
}

static float collect_ray( b2ShapeId shape, b2Vec2 point, b2Vec2 normal, float fraction, void* context )
{
	HitRec r = { shape.index1 - 1, point.x, point.y, normal.x, normal.y, fraction };
	hit_push( context, r );
	return 1.0f;
}

static bool collect_overlap( b2ShapeId shape, void* context )
{
	HitRec r = { shape.index1 - 1, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
	hit_push( context, r );
	return true;
}

// qtime: closest-ray queries of each world timed after `steps` steps (queries grouped by world)
typedef struct
{
	const Scene* sc;
	const QueryRec* qs;
	const long* start; // per world k of the sample: its first query; start[count] = the total
	int first, count, steps, next;
	double* world_ms;
	long* hits;
	pthread_mutex_t lock;
} QTimeJob;

static void* qtime_worker( void* arg )
{
	QTimeJob* job = arg;
	for ( ;; )
	{
		pthread_mutex_lock( &job->lock );
		int k = job->next++;
		pthread_mutex_unlock( &job->lock );
		if ( k >= job->count )
			break;
		const SceneWorld* sw = job->sc->worlds + job->first + k;
		b2BodyId* ids = malloc( sizeof( b2BodyId ) * ( sw->hdr.body_count + 1 ) );
		b2WorldId wid = build_world( sw, ids );
		for ( int s = 0; s < job->steps; ++s )
			b2World_Step( wid, 1.0f / 60.0f, 4 );
		long hits = 0;
		double t0 = now_sec();
		for ( long i = job->start[k]; i < job->start[k + 1]; ++i )
		{
			const QueryRec* q = job->qs + i;
			b2QueryFilter filter = b2DefaultQueryFilter();
			filter.categoryBits = (uint64_t)q->cat_lo | ( (uint64_t)q->cat_hi << 32 );
			filter.maskBits = (uint64_t)q->mask_lo | ( (uint64_t)q->mask_hi << 32 );
			b2RayResult r =
				b2World_CastRayClosest( wid, ( b2Vec2 ){ q->p[0], q->p[1] }, ( b2Vec2 ){ q->p[2], q->p[3] }, filter );
			hits += r.hit ? 1 : 0;
		}
		job->world_ms[k] = 1e3 * ( now_sec() - t0 );
		job->hits[k] = hits;
		pthread_mutex_lock( &g_world_lock );
		b2DestroyWorld( wid );
		pthread_mutex_unlock( &g_world_lock );
		free( ids );
	}
	return NULL;
}

// ---- API call fixtures ("ops" mode): one body, shape or joint call of the reference's API
// per record, applied before step `step` (tests/golden/make_golden.py OPS)
typedef struct
{
	int32_t step, op, target, flag;
	float f[6];
} OpRec; // 40 bytes

static uint32_t fbits( float v )
{
	uint32_t u;
	memcpy( &u, &v, 4 );
	return u;
}

static int g_jointCount; // joints of the scene plus the ones ops 25 created

static void apply_op( const OpRec* o, const b2BodyId* bodies, const b2ShapeId* shapes, b2JointId* joints )
{
	const float* f = o->f;
	b2BodyId b = bodies[o->target];
	switch ( o->op )
	{
		case 1:
			b2Body_ApplyForce( b, ( b2Vec2 ){ f[0], f[1] }, ( b2Vec2 ){ f[2], f[3] }, o->flag != 0 );
			break;
		case 2:
			b2Body_ApplyForceToCenter( b, ( b2Vec2 ){ f[0], f[1] }, o->flag != 0 );
			break;
		case 3:
			b2Body_ApplyTorque( b, f[0], o->flag != 0 );
			break;
		case 4:
			b2Body_ApplyLinearImpulse( b, ( b2Vec2 ){ f[0], f[1] }, ( b2Vec2 ){ f[2], f[3] }, o->flag != 0 );
			break;
		case 5:
			b2Body_ApplyLinearImpulseToCenter( b, ( b2Vec2 ){ f[0], f[1] }, o->flag != 0 );
			break;
		case 6:
			b2Body_ApplyAngularImpulse( b, f[0], o->flag != 0 );
			break;
		case 7:
			b2Body_SetTransform( b, ( b2Vec2 ){ f[0], f[1] }, ( b2Rot ){ f[2], f[3] } );
			break;
		case 8:
			b2Body_SetLinearVelocity( b, ( b2Vec2 ){ f[0], f[1] } );
			break;
		case 9:
			b2Body_SetAngularVelocity( b, f[0] );
			break;
		case 10:
		{
			b2MassData md = { f[0], { f[2], f[3] }, f[1] };
			b2Body_SetMassData( b, md );
		}
		break;
		case 11:
			b2Body_ApplyMassFromShapes( b );
			break;
		case 12:
			b2Shape_SetDensity( shapes[o->target], f[0], o->flag != 0 );
			break;
		case 13:
			b2Shape_SetFriction( shapes[o->target], f[0] );
			break;
		case 14:
			b2Shape_SetRestitution( shapes[o->target], f[0] );
			break;
		case 15:
			b2Body_SetGravityScale( b, f[0] );
			break;
		case 16:
			b2Body_SetLinearDamping( b, f[0] );
			break;
		case 17:
			b2Body_SetAngularDamping( b, f[0] );
			break;
		case 18:
			b2Body_SetFixedRotation( b, o->flag != 0 );
			break;
		case 19:
			b2Body_SetBullet( b, o->flag != 0 );
			break;
		case 20:
			b2Body_SetAwake( b, o->flag != 0 );
			break;
		case 21:
		{
			b2Filter filter = { (uint64_t)fbits( f[0] ) | ( (uint64_t)fbits( f[1] ) << 32 ),
								(uint64_t)fbits( f[2] ) | ( (uint64_t)fbits( f[3] ) << 32 ), (int)fbits( f[4] ) };
			b2Shape_SetFilter( shapes[o->target], filter );
		}
		break;
		case 22:
			b2DestroyJoint( joints[o->target] );
			break;
		case 23:
			b2Joint_WakeBodies( joints[o->target] );
			break;
		case 24:
			b2Joint_SetCollideConnected( joints[o->target], o->flag != 0 );
			break;
		case 25:
		{
			// a prismatic joint between bodies target and (flag & 0xFFFF), anchors f0..f3,
			// axis f4, f5, collideConnected flag >> 16, everything else the default def
			b2PrismaticJointDef def = b2DefaultPrismaticJointDef();
			def.bodyIdA = b;
			def.bodyIdB = bodies[o->flag & 0xFFFF];
			def.localAnchorA = ( b2Vec2 ){ f[0], f[1] };
			def.localAnchorB = ( b2Vec2 ){ f[2], f[3] };
			def.localAxisA = ( b2Vec2 ){ f[4], f[5] };
			def.collideConnected = ( o->flag >> 16 ) != 0;
			joints[g_jointCount++] = b2CreatePrismaticJoint( b2Body_GetWorld( b ), &def );
		}
		break;
		case 26:
		{
			// prismatic joint target: limits f0, f1, motor force f2 and speed f3, spring
			// hertz f4 and damping f5, then limit / motor / spring on per flag bits 1, 2, 4
			b2JointId j = joints[o->target];
			b2PrismaticJoint_SetLimits( j, f[0], f[1] );
			b2PrismaticJoint_EnableLimit( j, ( o->flag & 1 ) != 0 );
			b2PrismaticJoint_SetMaxMotorForce( j, f[2] );
			b2PrismaticJoint_SetMotorSpeed( j, f[3] );
			b2PrismaticJoint_EnableMotor( j, ( o->flag & 2 ) != 0 );
			b2PrismaticJoint_SetSpringHertz( j, f[4] );
			b2PrismaticJoint_SetSpringDampingRatio( j, f[5] );
			b2PrismaticJoint_EnableSpring( j, ( o->flag & 4 ) != 0 );
		}
		break;
		case 27:
			b2PrismaticJoint_SetTargetTranslation( joints[o->target], f[0] );
			break;
		case 28:
		{
			// a weld joint between bodies target and (flag & 0xFFFF), anchors f0..f3,
			// reference angle f4, collideConnected flag >> 16, stiffness from the default def
			b2WeldJointDef def = b2DefaultWeldJointDef();
			def.bodyIdA = b;
			def.bodyIdB = bodies[o->flag & 0xFFFF];
			def.localAnchorA = ( b2Vec2 ){ f[0], f[1] };
			def.localAnchorB = ( b2Vec2 ){ f[2], f[3] };
			def.referenceAngle = f[4];
			def.collideConnected = ( o->flag >> 16 ) != 0;
			joints[g_jointCount++] = b2CreateWeldJoint( b2Body_GetWorld( b ), &def );
		}
		break;
		case 29:
		{
			// weld joint target: linear hertz f0, angular hertz f1, linear damping f2, angular damping f3
			b2JointId j = joints[o->target];
			b2WeldJoint_SetLinearHertz( j, f[0] );
			b2WeldJoint_SetAngularHertz( j, f[1] );
			b2WeldJoint_SetLinearDampingRatio( j, f[2] );
			b2WeldJoint_SetAngularDampingRatio( j, f[3] );
		}
		break;
		case 30:
		{
			// a wheel joint from the default def (spring on, 1 Hz, damping 0.7), anchors
			// f0..f3, axis f4, f5, collideConnected flag >> 16
			b2WheelJointDef def = b2DefaultWheelJointDef();
			def.bodyIdA = b;
			def.bodyIdB = bodies[o->flag & 0xFFFF];
			def.localAnchorA = ( b2Vec2 ){ f[0], f[1] };
			def.localAnchorB = ( b2Vec2 ){ f[2], f[3] };
			def.localAxisA = ( b2Vec2 ){ f[4], f[5] };
			def.collideConnected = ( o->flag >> 16 ) != 0;
			joints[g_jointCount++] = b2CreateWheelJoint( b2Body_GetWorld( b ), &def );
		}
		break;
		case 31:
		{
			// wheel joint target: limits f0, f1, motor torque f2 and speed f3, spring hertz
			// f4 and damping f5, then limit / motor / spring on per flag bits 1, 2, 4
			b2JointId j = joints[o->target];
			b2WheelJoint_SetLimits( j, f[0], f[1] );
			b2WheelJoint_EnableLimit( j, ( o->flag & 1 ) != 0 );
			b2WheelJoint_SetMaxMotorTorque( j, f[2] );
			b2WheelJoint_SetMotorSpeed( j, f[3] );
			b2WheelJoint_EnableMotor( j, ( o->flag & 2 ) != 0 );
			b2WheelJoint_SetSpringHertz( j, f[4] );
			b2WheelJoint_SetSpringDampingRatio( j, f[5] );
			b2WheelJoint_EnableSpring( j, ( o->flag & 4 ) != 0 );
		}
		break;
		case 32:
		{
			// a distance joint from the default def, anchors f0..f3, length f4,
			// collideConnected flag >> 16
			b2DistanceJointDef def = b2DefaultDistanceJointDef();
			def.bodyIdA = b;
			def.bodyIdB = bodies[o->flag & 0xFFFF];
			def.localAnchorA = ( b2Vec2 ){ f[0], f[1] };
			def.localAnchorB = ( b2Vec2 ){ f[2], f[3] };
			def.length = f[4];
			def.collideConnected = ( o->flag >> 16 ) != 0;
			joints[g_jointCount++] = b2CreateDistanceJoint( b2Body_GetWorld( b ), &def );
		}
		break;
		case 33:
		{
			// distance joint target: length range f0, f1, motor force f2 and speed f3,
			// spring hertz f4 and damping f5, then limit / motor / spring per flag bits 1, 2, 4
			b2JointId j = joints[o->target];
			b2DistanceJoint_SetLengthRange( j, f[0], f[1] );
			b2DistanceJoint_EnableLimit( j, ( o->flag & 1 ) != 0 );
			b2DistanceJoint_SetMaxMotorForce( j, f[2] );
			b2DistanceJoint_SetMotorSpeed( j, f[3] );
			b2DistanceJoint_EnableMotor( j, ( o->flag & 2 ) != 0 );
			b2DistanceJoint_SetSpringHertz( j, f[4] );
			b2DistanceJoint_SetSpringDampingRatio( j, f[5] );
			b2DistanceJoint_EnableSpring( j, ( o->flag & 4 ) != 0 );
		}
		break;
		case 34:
			b2DistanceJoint_SetLength( joints[o->target], f[0] );
			break;
		case 35:
		{
			// revolute joint target: limits f0, f1, motor torque f2 and speed f3, spring
			// hertz f4 and damping f5, then limit / motor / spring per flag bits 1, 2, 4
			b2JointId j = joints[o->target];
			b2RevoluteJoint_SetLimits( j, f[0], f[1] );
			b2RevoluteJoint_EnableLimit( j, ( o->flag & 1 ) != 0 );
			b2RevoluteJoint_SetMaxMotorTorque( j, f[2] );
			b2RevoluteJoint_SetMotorSpeed( j, f[3] );
			b2RevoluteJoint_EnableMotor( j, ( o->flag & 2 ) != 0 );
			b2RevoluteJoint_SetSpringHertz( j, f[4] );
			b2RevoluteJoint_SetSpringDampingRatio( j, f[5] );
			b2RevoluteJoint_EnableSpring( j, ( o->flag & 4 ) != 0 );
		}
		break;
		case 36:
			b2RevoluteJoint_SetTargetAngle( joints[o->target], f[0] );
			break;
		case 37:
		{
			// a motor joint: linear offset f0, f1, angular offset f2, maximum force f3 and
			// torque f4, correction factor f5, collideConnected flag >> 16
			b2MotorJointDef def = b2DefaultMotorJointDef();
			def.bodyIdA = b;
			def.bodyIdB = bodies[o->flag & 0xFFFF];
			def.linearOffset = ( b2Vec2 ){ f[0], f[1] };
			def.angularOffset = f[2];
			def.maxForce = f[3];
			def.maxTorque = f[4];
			def.correctionFactor = f[5];
			def.collideConnected = ( o->flag >> 16 ) != 0;
			joints[g_jointCount++] = b2CreateMotorJoint( b2Body_GetWorld( b ), &def );
		}
		break;
		case 38:
		{
			b2JointId j = joints[o->target];
			b2MotorJoint_SetLinearOffset( j, ( b2Vec2 ){ f[0], f[1] } );
			b2MotorJoint_SetAngularOffset( j, f[2] );
			b2MotorJoint_SetMaxForce( j, f[3] );
			b2MotorJoint_SetMaxTorque( j, f[4] );
			b2MotorJoint_SetCorrectionFactor( j, f[5] );
		}
		break;
		case 39:
		{
			// a mouse joint: target f0, f1, hertz f2, damping f3, maximum force f4,
			// collideConnected flag >> 16
			b2MouseJointDef def = b2DefaultMouseJointDef();
			def.bodyIdA = b;
			def.bodyIdB = bodies[o->flag & 0xFFFF];
			def.target = ( b2Vec2 ){ f[0], f[1] };
			def.hertz = f[2];
			def.dampingRatio = f[3];
			def.maxForce = f[4];
			def.collideConnected = ( o->flag >> 16 ) != 0;
			joints[g_jointCount++] = b2CreateMouseJoint( b2Body_GetWorld( b ), &def );
		}
		break;
		case 40:
		{
			b2JointId j = joints[o->target];
			b2MouseJoint_SetTarget( j, ( b2Vec2 ){ f[0], f[1] } );
			b2MouseJoint_SetSpringHertz( j, f[2] );
			b2MouseJoint_SetSpringDampingRatio( j, f[3] );
			b2MouseJoint_SetMaxForce( j, f[4] );
		}
		break;
		case 42:
			b2DestroyJoint( joints[o->target] );
			break;
		case 43:
			b2Joint_SetCollideConnected( joints[o->target], o->flag != 0 );
			break;
		case 41:
		{
			b2FilterJointDef def = b2DefaultFilterJointDef();
			def.bodyIdA = b;
			def.bodyIdB = bodies[o->flag & 0xFFFF];
			joints[g_jointCount++] = b2CreateFilterJoint( b2Body_GetWorld( b ), &def );
		}
		break;
		default:
			fprintf( stderr, "bad op %d\n", o->op );
			exit( 3 );
	}
}

int main( int argc, char** argv )
{
	if ( argc < 3 )
	{
		fprintf( stderr, "usage: ref_harness dump|time|pairs ...\n" );
		return 2;
	}
	if ( strcmp( argv[1], "dump" ) == 0 && argc >= 7 )
	{
		Scene sc = load_scene( argv[2] );
		const char* out = argv[3];
		int wi = atoi( argv[4] );
		int steps = atoi( argv[5] );
		int dumps[4096];
		int nd = parse_list( argv[6], dumps, 4096 );
		int substeps = argc >= 8 ? atoi( argv[7] ) : 4;
		const SceneWorld* sw = sc.worlds + wi;
		b2BodyId* ids = malloc( sizeof( b2BodyId ) * ( sw->hdr.body_count + 1 ) );
		b2WorldId wid = build_world( sw, ids );
		FILE* f = fopen( out, "wb" );
		int di = 0;
		if ( di < nd && dumps[di] == 0 )
		{
			write_snapshot( f, wid, 0 );
			di++;
		}
		for ( int s = 1; s <= steps; ++s )
		{
			b2World_Step( wid, 1.0f / 60.0f, substeps );
			while ( di < nd && dumps[di] == s )
			{
				write_snapshot( f, wid, s );
				di++;
			}
		}
		fclose( f );
		b2DestroyWorld( wid );
		return 0;
	}
	if ( strcmp( argv[1], "ops" ) == 0 && argc >= 8 )
	{
		// ref_harness ops <scene> <ops> <out> <world> <steps> <dumps>: the world stepped
		// like "dump", with body, shape and joint API calls applied before the steps
		// the ops file names (tests/golden/make_golden.py writes it)
		Scene sc = load_scene( argv[2] );
		FILE* of = fopen( argv[3], "rb" );
		if ( !of )
			return 2;
		fseek( of, 0, SEEK_END );
		long ob = ftell( of );
		fseek( of, 0, SEEK_SET );
		int nops = (int)( ob / sizeof( OpRec ) );
		OpRec* ops = malloc( sizeof( OpRec ) * ( nops + 1 ) );
		if ( fread( ops, sizeof( OpRec ), nops, of ) != (size_t)nops )
			return 2;
		fclose( of );
		const char* out = argv[4];
		int wi = atoi( argv[5] );
		int steps = atoi( argv[6] );
		int dumps[4096];
		int nd = parse_list( argv[7], dumps, 4096 );
		const SceneWorld* sw = sc.worlds + wi;
		b2BodyId* ids = malloc( sizeof( b2BodyId ) * ( sw->hdr.body_count + 1 ) );
		g_shapeIds = malloc( sizeof( b2ShapeId ) * ( sw->hdr.shape_count + 1 ) );
		g_jointIds = malloc( sizeof( b2JointId ) * ( sw->hdr.joint_count + nops + 1 ) );
		g_jointCount = sw->hdr.joint_count;
		b2WorldId wid = build_world( sw, ids );
		FILE* f = fopen( out, "wb" );
		int di = 0, oi = 0;
		if ( di < nd && dumps[di] == 0 )
		{
			write_snapshot( f, wid, 0 );
			di++;
		}
		for ( int s = 1; s <= steps; ++s )
		{
			while ( oi < nops && ops[oi].step == s )
			{
				apply_op( ops + oi, ids, g_shapeIds, g_jointIds );
				oi++;
			}
			b2World_Step( wid, 1.0f / 60.0f, 4 );
			while ( di < nd && dumps[di] == s )
			{
				write_snapshot( f, wid, s );
				di++;
			}
		}
		fclose( f );
		b2DestroyWorld( wid );
		return 0;
	}
	if ( strcmp( argv[1], "geom" ) == 0 && argc >= 5 )
	{
		// per shape, 148 bytes: int32 type, count; float radius, centroid x/y,
		// 8 vertices, 8 normals (circle: v0 = center; capsule/segment: v0, v1)
		Scene sc = load_scene( argv[2] );
		int wi = atoi( argv[4] );
		const SceneWorld* sw = sc.worlds + wi;
		b2BodyId* ids = malloc( sizeof( b2BodyId ) * ( sw->hdr.body_count + 1 ) );
		b2WorldId wid = build_world( sw, ids );
		b2World* world = b2GetWorldFromId( wid );
		FILE* f = fopen( argv[3], "wb" );
		for ( int i = 0; i < world->shapes.count; ++i )
		{
			b2Shape* shape = world->shapes.data + i;
			int32_t head[2] = { (int32_t)shape->type, 0 };
			float rec[35];
			memset( rec, 0, sizeof( rec ) );
			switch ( shape->type )
			{
				case b2_circleShape:
					head[1] = 1;
					rec[0] = shape->circle.radius;
					rec[1] = shape->circle.center.x;
					rec[2] = shape->circle.center.y;
					rec[3] = shape->circle.center.x;
					rec[4] = shape->circle.center.y;
					break;
				case b2_capsuleShape:
					head[1] = 2;
					rec[0] = shape->capsule.radius;
					rec[3] = shape->capsule.center1.x;
					rec[4] = shape->capsule.center1.y;
					rec[5] = shape->capsule.center2.x;
					rec[6] = shape->capsule.center2.y;
					break;
				case b2_segmentShape:
					head[1] = 2;
					rec[3] = shape->segment.point1.x;
					rec[4] = shape->segment.point1.y;
					rec[5] = shape->segment.point2.x;
					rec[6] = shape->segment.point2.y;
					break;
				case b2_polygonShape:
					head[1] = shape->polygon.count;
					rec[0] = shape->polygon.radius;
					rec[1] = shape->polygon.centroid.x;
					rec[2] = shape->polygon.centroid.y;
					for ( int k = 0; k < shape->polygon.count; ++k )
					{
						rec[3 + 2 * k] = shape->polygon.vertices[k].x;
						rec[4 + 2 * k] = shape->polygon.vertices[k].y;
						rec[19 + 2 * k] = shape->polygon.normals[k].x;
						rec[20 + 2 * k] = shape->polygon.normals[k].y;
					}
					break;
				default:
					break;
			}
			fwrite( head, sizeof( head ), 1, f );
			fwrite( rec, sizeof( rec ), 1, f );
		}
		fclose( f );
		b2DestroyWorld( wid );
		return 0;
	}
	if ( strcmp( argv[1], "events" ) == 0 && argc >= 7 )
	{
		// per listed step: int32 [step, begin, end, hit, move] counts, then float64 records
		// in the layout of the flat shim's event getters (F/native/wasm/f2d_shim.c:1186-1300):
		// begin 13 [aI1, 0, bI1, 0, nx, ny, pointCount, p0x, p0y, sep0, p1x, p1y, sep1],
		// end 4 [aI1, 0, bI1, 0], hit 9 [aI1, 0, bI1, 0, px, py, nx, ny, approachSpeed],
		// move 7 [bodyI1, 0, x, y, qc, qs, fellAsleep]
		Scene sc = load_scene( argv[2] );
		int wi = atoi( argv[4] );
		int steps = atoi( argv[5] );
		int dumps[4096];
		int nd = parse_list( argv[6], dumps, 4096 );
		const SceneWorld* sw = sc.worlds + wi;
		b2BodyId* ids = malloc( sizeof( b2BodyId ) * ( sw->hdr.body_count + 1 ) );
		b2WorldId wid = build_world( sw, ids );
		FILE* f = fopen( argv[3], "wb" );
		int di = 0;
		for ( int s = 1; s <= steps; ++s )
		{
			b2World_Step( wid, 1.0f / 60.0f, 4 );
			if ( di >= nd || dumps[di] != s )
			{
				continue;
			}
			di++;
			b2ContactEvents ce = b2World_GetContactEvents( wid );
			b2BodyEvents be = b2World_GetBodyEvents( wid );
			int32_t counts[5] = { s, ce.beginCount, ce.endCount, ce.hitCount, be.moveCount };
			fwrite( counts, sizeof( counts ), 1, f );
			for ( int i = 0; i < ce.beginCount; ++i )
			{
				const b2ContactBeginTouchEvent* e = ce.beginEvents + i;
				double r[13] = { e->shapeIdA.index1, 0, e->shapeIdB.index1, 0, e->manifold.normal.x, e->manifold.normal.y,
								 e->manifold.pointCount };
				for ( int p = 0; p < 2; ++p )
				{
					r[7 + 3 * p] = e->manifold.points[p].point.x;
					r[8 + 3 * p] = e->manifold.points[p].point.y;
					r[9 + 3 * p] = e->manifold.points[p].separation;
				}
				fwrite( r, sizeof( r ), 1, f );
			}
			for ( int i = 0; i < ce.endCount; ++i )
			{
				const b2ContactEndTouchEvent* e = ce.endEvents + i;
				double r[4] = { e->shapeIdA.index1, 0, e->shapeIdB.index1, 0 };
				fwrite( r, sizeof( r ), 1, f );
			}
			for ( int i = 0; i < ce.hitCount; ++i )
			{
				const b2ContactHitEvent* e = ce.hitEvents + i;
				double r[9] = { e->shapeIdA.index1, 0, e->shapeIdB.index1, 0, e->point.x, e->point.y, e->normal.x,
								e->normal.y, e->approachSpeed };
				fwrite( r, sizeof( r ), 1, f );
			}
			for ( int i = 0; i < be.moveCount; ++i )
			{
				const b2BodyMoveEvent* e = be.moveEvents + i;
				double r[7] = { e->bodyId.index1, 0, e->transform.p.x, e->transform.p.y, e->transform.q.c, e->transform.q.s,
								e->fellAsleep };
				fwrite( r, sizeof( r ), 1, f );
			}
		}
		fclose( f );
		b2DestroyWorld( wid );
		return 0;
	}
	if ( strcmp( argv[1], "query" ) == 0 && argc >= 7 )
	{
		// <out>: per query int32 count, then count HitRec in the reference's callback
		// order (closest ray: 0 or 1; every ray hit; overlapping shapes; explosion: 0).
		// <out>.snap: the world snapshot after the steps, then after the queries.
		Scene sc = load_scene( argv[2] );
		FILE* qf = fopen( argv[3], "rb" );
		if ( !qf )
		{
			return 1;
		}
		fseek( qf, 0, SEEK_END );
		long qbytes = ftell( qf );
		fseek( qf, 0, SEEK_SET );
		int nq = (int)( qbytes / (long)sizeof( QueryRec ) );
		QueryRec* qs = malloc( sizeof( QueryRec ) * ( nq + 1 ) );
		if ( fread( qs, sizeof( QueryRec ), nq, qf ) != (size_t)nq )
		{
			return 1;
		}
		fclose( qf );
		int wi = atoi( argv[5] );
		int steps = atoi( argv[6] );
		const SceneWorld* sw = sc.worlds + wi;
		b2BodyId* ids = malloc( sizeof( b2BodyId ) * ( sw->hdr.body_count + 1 ) );
		b2WorldId wid = build_world( sw, ids );
		for ( int s = 1; s <= steps; ++s )
		{
			b2World_Step( wid, 1.0f / 60.0f, 4 );
		}
		char snapPath[4096];
		snprintf( snapPath, sizeof( snapPath ), "%s.snap", argv[4] );
		FILE* sf = fopen( snapPath, "wb" );
		write_snapshot( sf, wid, steps );
		FILE* f = fopen( argv[4], "wb" );
		for ( int i = 0; i < nq; ++i )
		{
			const QueryRec* q = qs + i;
			b2QueryFilter filter = b2DefaultQueryFilter();
			filter.categoryBits = (uint64_t)q->cat_lo | ( (uint64_t)q->cat_hi << 32 );
			filter.maskBits = (uint64_t)q->mask_lo | ( (uint64_t)q->mask_hi << 32 );
			HitList list = { 0 };
			if ( q->kind == 1 )
			{
				b2RayResult r = b2World_CastRayClosest( wid, ( b2Vec2 ){ q->p[0], q->p[1] }, ( b2Vec2 ){ q->p[2], q->p[3] },
														filter );
				if ( r.hit )
				{
					HitRec h = { r.shapeId.index1 - 1, r.point.x, r.point.y, r.normal.x, r.normal.y, r.fraction };
					hit_push( &list, h );
				}
			}
			else if ( q->kind == 2 )
			{
				b2World_CastRay( wid, ( b2Vec2 ){ q->p[0], q->p[1] }, ( b2Vec2 ){ q->p[2], q->p[3] }, filter, collect_ray,
								 &list );
			}
			else if ( q->kind == 3 )
			{
				b2AABB box = { { q->p[0], q->p[1] }, { q->p[2], q->p[3] } };
				b2World_OverlapAABB( wid, box, filter, collect_overlap, &list );
			}
			else if ( q->kind == 4 )
			{
				b2ExplosionDef def = b2DefaultExplosionDef();
				def.maskBits = (uint64_t)q->mask_lo | ( (uint64_t)q->mask_hi << 32 );
				def.position = ( b2Vec2 ){ q->p[0], q->p[1] };
				def.radius = q->p[2];
				def.falloff = q->p[3];
				def.impulsePerLength = q->p[4];
				b2World_Explode( wid, &def );
			}
			int32_t count = list.count;
			fwrite( &count, sizeof( count ), 1, f );
			if ( list.count )
			{
				fwrite( list.hits, sizeof( HitRec ), list.count, f );
			}
			free( list.hits );
		}
		fclose( f );
		write_snapshot( sf, wid, steps );
		fclose( sf );
		b2DestroyWorld( wid );
		return 0;
	}
	if ( strcmp( argv[1], "qtime" ) == 0 && argc >= 8 )
	{
		// ref_harness qtime <scene> <queries> <steps> <threads> <first_world> <count>
		Scene sc = load_scene( argv[2] );
		FILE* qf = fopen( argv[3], "rb" );
		if ( !qf )
		{
			return 1;
		}
		fseek( qf, 0, SEEK_END );
		long nq = ftell( qf ) / (long)sizeof( QueryRec );
		fseek( qf, 0, SEEK_SET );
		QueryRec* qs = malloc( sizeof( QueryRec ) * ( nq + 1 ) );
		if ( fread( qs, sizeof( QueryRec ), nq, qf ) != (size_t)nq )
		{
			return 1;
		}
		fclose( qf );
		QTimeJob job;
		memset( &job, 0, sizeof( job ) );
		job.sc = &sc;
		job.qs = qs;
		job.steps = atoi( argv[4] );
		int threads = atoi( argv[5] );
		job.first = atoi( argv[6] );
		job.count = atoi( argv[7] );
		long* start = calloc( (size_t)job.count + 1, sizeof( long ) );
		long i = 0;
		for ( int k = 0; k <= job.count; ++k )
		{
			while ( i < nq && qs[i].world < job.first + k )
				++i;
			start[k] = i;
		}
		job.start = start;
		job.world_ms = calloc( (size_t)job.count, sizeof( double ) );
		job.hits = calloc( (size_t)job.count, sizeof( long ) );
		pthread_mutex_init( &job.lock, NULL );
		pthread_t tid[256];
		if ( threads > 256 )
			threads = 256;
		double t0 = now_sec();
		for ( int t = 0; t < threads; ++t )
			pthread_create( tid + t, NULL, qtime_worker, &job );
		for ( int t = 0; t < threads; ++t )
			pthread_join( tid[t], NULL );
		double wall = now_sec() - t0;
		double total_ms = 0;
		long hits = 0;
		for ( int k = 0; k < job.count; ++k )
		{
			total_ms += job.world_ms[k];
			hits += job.hits[k];
		}
		long rays = start[job.count] - start[0];
		int used = threads < job.count ? threads : job.count;
		double busy = total_ms / 1e3 / ( used > 0 ? used : 1 );
		printf( "{\"worlds\": %d, \"rays\": %ld, \"hits\": %ld, \"steps\": %d, \"threads\": %d, \"wall_s\": %.4f, "
				"\"query_s\": %.6f, \"rays_per_s\": %.6e}\n",
				job.count, rays, hits, job.steps, used, wall, busy, rays / ( busy > 0 ? busy : 1e-9 ) );
		return 0;
	}
	if ( strcmp( argv[1], "time" ) == 0 && argc >= 6 )
	{
		Scene sc = load_scene( argv[2] );
		TimeJob job;
		memset( &job, 0, sizeof( job ) );
		job.sc = &sc;
		job.steps = atoi( argv[3] );
		job.warmup = atoi( argv[4] );
		int threads = atoi( argv[5] );
		job.first = argc >= 7 ? atoi( argv[6] ) : 0;
		job.count = argc >= 8 ? atoi( argv[7] ) : sc.count - job.first;
		job.substeps = 4;
		pthread_mutex_init( &job.lock, NULL );
		job.step_ms = calloc( (size_t)job.count * job.steps, sizeof( double ) );
		double bodies = 0;
		for ( int k = 0; k < job.count; ++k )
		{
			const SceneWorld* sw = sc.worlds + job.first + k;
			for ( int b = 0; b < sw->hdr.body_count; ++b )
				bodies += sw->bodies[b].type != 0 ? 1 : 0;
		}
		pthread_t tid[256];
		double t0 = now_sec();
		for ( int t = 0; t < threads; ++t )
			pthread_create( tid + t, NULL, time_worker, &job );
		for ( int t = 0; t < threads; ++t )
			pthread_join( tid[t], NULL );
		double wall = now_sec() - t0;
		double total_ms = 0;
		long n = (long)job.count * job.steps;
		for ( long i = 0; i < n; ++i )
			total_ms += job.step_ms[i];
		qsort( job.step_ms, n, sizeof( double ), cmp_double );
		double p50 = job.step_ms[n / 2];
		double p99 = job.step_ms[(long)( 0.99 * ( n - 1 ) )];
		double body_substeps = bodies * job.steps * job.substeps;
		// throughput over the timed steps only: per-thread time is summed step time,
		// over the threads that had a world to step
		int used = threads < job.count ? threads : job.count;
		double busy = total_ms / 1e3 / ( used > 0 ? used : 1 );
		printf( "{\"worlds\": %d, \"bodies\": %.0f, \"steps\": %d, \"warmup\": %d, \"threads\": %d, "
				"\"wall_s\": %.4f, \"step_ms_mean\": %.6f, \"step_ms_p50\": %.6f, \"step_ms_p99\": %.6f, "
				"\"body_substeps_per_s\": %.6e}\n",
				job.count, bodies, job.steps, job.warmup, used, wall, total_ms / n, p50, p99, body_substeps / busy );
		return 0;
	}
	fprintf( stderr, "bad arguments\n" );
	return 2;
}
